"""CPU oracle for the two-site effective-Hamiltonian path (TEST INFRASTRUCTURE ONLY).

This package is the oracle the tests check against, never the product. Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import it. The product
package `tnkit_amd` never imports it and has no CPU fallback: its kernels run through
`libtnkit_hip.so` or fail.

Contents
--------
dense      numpy float64/float32 restatement of tnkit's dense two-site DMRG path: the
           contractions, environment updates, thick-restart / mixed Lanczos, the fixed-m
           Krylov exponential, SVD and Gram truncation, `random_mps`, `right_canonicalize`
           and the sweep. Each function cites the tnkit file:line it restates.
exact      closed forms and brute force that pin `dense`: the free-fermion TFIM energy
           (uniform and per-site fields), dense Hamiltonians from an MPO, exact
           diagonalisation of small chains.

How it is pinned (each function is checked by something that is not a retyping of it):
  * contractions / env: a single brute-force `np.einsum` over the whole network, the
    Hermiticity invariant <x, H y> = <H x, y>, the dense H built from the MPO for tiny chains,
    and the golden vectors written from tnkit itself (tests/golden/).
  * Lanczos: numpy.linalg.eigh on random symmetric matrices, breakdown on an exact eigenvector,
    the restart boundary; the golden (eigenvalue, iters, restarts) triples from tnkit.
  * Krylov exponential: scipy.linalg.expm.
  * truncations: numpy.linalg.svd (isometry, discarded weight, Eckart-Young optimality).
  * DMRG: the exact free-fermion energy of the open TFIM chain, exact diagonalisation for
    small chains, and tnkit's own per-sweep energies / matvec counts (golden fixtures).
"""
