"""Solver modules (spectralDNS/solvers); the HIP path serves NS."""
