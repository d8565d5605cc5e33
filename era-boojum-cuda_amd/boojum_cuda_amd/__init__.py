"""MI355X (gfx950) batched Goldilocks NTT / coset-LDE, a drop-in for
boojum-cuda's `ntt` and `context` modules (`ntt`, `context`), the order
conversion and domain powers beside them (`ops`), barycentric evaluation
(`barycentric`) and Poseidon Merkle-tree hashing (`poseidon`), with the LDE, ragged-batch and serving extensions (`lde`,
`ragged`, `serve`, `autotune`).  Compute runs in the HIP library
libboojum_ntt.so (C ABI: include/boojum_ntt.h)."""
from . import barycentric, context, lde, ntt, ops, poseidon, ragged, serve  # noqa: F401
from ._lib import HipError, LIB_PATH  # noqa: F401

P = 0xFFFFFFFF00000001
