"""sdns_spectrum_noalloc (partial sums in the library's static device scratch)
against sdns_spectrum (partial sums in the plan's workspace), through the
C-ABI: two calls in a row, the second on a smaller box with fewer shells; a
plan whose workspace was never allocated stays without one; nbins beyond the
scratch's limit is rejected.

Tolerance: the shell sums are FP64 sums of the same terms in a different
order (other block counts, atomic adds in LDS), so they agree to 1e-13
relative; the mode counts are exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
TWO_PI = 2 * np.pi


def _field(N, precision, seed):
    from spectraldns_amd.isotropic import random_solenoidal_field
    u = random_solenoidal_field(N, (TWO_PI,) * 3, seed=seed, k0=4.0, energy=0.5)
    cd = torch.complex128 if precision == "double" else torch.complex64
    return torch.from_numpy(u).to(cd).cuda()


def _compare(plan, u, nbins):
    out_ref = torch.empty(2 * nbins, dtype=torch.float64, device="cuda")
    out_na = torch.full((2 * nbins,), np.nan, dtype=torch.float64, device="cuda")
    plan.spectrum_sums(u, nbins, out=out_ref)
    plan.spectrum_sums(u, nbins, out=out_na, noalloc=True)
    torch.cuda.synchronize()
    ref, na = out_ref.cpu().numpy(), out_na.cpu().numpy()
    assert np.all(np.isfinite(na))
    np.testing.assert_array_equal(na[nbins:], ref[nbins:])              # mode counts
    scale = max(float(np.abs(ref[:nbins]).max()), 1e-300)
    assert float(np.abs(na[:nbins] - ref[:nbins]).max()) <= 1e-13 * scale
    assert ref[nbins:].sum() > 0
    return na


def test_spectrum_noalloc_matches_two_calls():
    from spectraldns_amd._abi import Plan
    print("\n[case] spectrum_noalloc 48^3 FP64 (40 shells), then 16^3 FP32 (10 shells)", flush=True)
    p1 = Plan((48, 48, 48), (TWO_PI,) * 3, "double", "2/3-rule")
    first = _compare(p1, _field((48, 48, 48), "double", 7), 40)
    p2 = Plan((16, 16, 16), (TWO_PI,) * 3, "single", "3/2-rule")
    second = _compare(p2, _field((16, 16, 16), "single", 8), 10)
    # the second call's output is its own, not a leftover of the first's scratch
    assert not np.array_equal(first[:10], second[:10])


def test_spectrum_noalloc_leaves_workspace_unallocated():
    from spectraldns_amd._abi import Plan, SdnsError
    print("\n[case] spectrum_noalloc on a fresh 128^3 FP64 plan (no workspace)", flush=True)
    N = (128, 128, 128)
    u = _field(N, "double", 9)
    out = torch.empty(2 * 120, dtype=torch.float64, device="cuda")
    p = Plan(N, (TWO_PI,) * 3, "double", "2/3-rule", tuned=False)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    p.spectrum_sums(u, 120, out=out, noalloc=True)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    ws = p._lib.sdns_workspace_bytes(p._h)
    assert ws > 64 << 20
    assert free0 - free1 < 16 << 20, (free0, free1, ws)   # the workspace was not allocated
    counts = out[120:].cpu().numpy()
    assert counts.sum() > 0
    with pytest.raises(SdnsError, match="no-alloc scratch"):
        p.spectrum_sums(u, 8193, out=torch.empty(2 * 8193, dtype=torch.float64, device="cuda"), noalloc=True)
