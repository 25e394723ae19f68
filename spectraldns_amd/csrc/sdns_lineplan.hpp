// Radix sequence of the line FFTs: shared by the device code and by the host
// code that sizes the batched kernels' tiles.
#pragma once

namespace sdns {

// Radix of pass `pass` for a length-M line with E points per thread: the
// largest divisor of E that divides what is left of M.
constexpr int radix_at(int M, int E, int pass) {
  int rem = M, r = 1;
  for (int p = 0; p <= pass; ++p) {
    if (rem == 1) return 1;
    r = 1;
    for (int d = E; d >= 2; --d)
      if (E % d == 0 && rem % d == 0) { r = d; break; }
    if (r == 1) return 0;  // not factorable with divisors of E
    rem /= r;
  }
  return r;
}
constexpr int num_passes(int M, int E) {
  int n = 0, rem = M;
  while (rem > 1) {
    int r = radix_at(M, E, n);
    if (r <= 1) return -1;
    rem /= r;
    ++n;
  }
  return n;
}

}  // namespace sdns
