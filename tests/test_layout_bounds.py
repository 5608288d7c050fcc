"""CPU checks of the population layout's index arithmetic: the host
instantiation of the functions every fluid and conversion kernel indexes
with (pop_index, tile_origin, xcd_remap, shift_axis), walked along each
kernel's index path, must keep every population index below 19 N (the
elements of one lattice in any storage) and every link-table index (the
wing's and the exact-root pass's) below tiles x 19 x 256, and the block ->
tile remap must be a bijection.  The exact-root pass's puller enumeration
(bf_rays.h) is walked both ways: each puller of a source pulls from it
(periodic wrap, outlet clamps), and each cell is among its source's
pullers.

Shapes: the four workloads (64^3 periodic shear wave, the 256x128x128
plate with an inlet / outlet pair, the 512x256x256 bird, the 1024x512x512
precision case), grids whose sides are not multiples of the 64 x 4 tile,
and the largest grid a 32-bit cell count allows along x.  Every face
treatment is walked: periodic, far field and outlet.  The storage only
scales the element size, so the element bound covers FP32, FP16 and BF16
(bytes = index x 4 or 2 < the allocation's 19 N x 4 or 2)."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "numi-gpu-physics_amd", "birdflow_hip", "libbirdflow_adjoint_check.so")
PERIODIC, FARFIELD, OUTLET = 0, 1, 2


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("check library not built")
    l = ctypes.CDLL(LIB)
    l.bfa_layout_check.argtypes = [ctypes.c_uint32] * 3 + [ctypes.POINTER(ctypes.c_int),
                                                          ctypes.POINTER(ctypes.c_uint64)]
    l.bfa_layout_check.restype = ctypes.c_int
    return l


def check(lib, shape, bc):
    flat = (ctypes.c_int * 6)(*[v for pair in bc for v in pair])
    out = (ctypes.c_uint64 * 5)()
    rc = lib.bfa_layout_check(*shape, flat, out)
    return rc, list(out)


SHAPES = [(64, 64, 64), (256, 128, 128), (512, 256, 256), (1024, 512, 512), (59, 53, 50), (100, 7, 3),
          (1, 1, 1), (65, 5, 2)]
BCS = {"periodic": ((PERIODIC, PERIODIC),) * 3,
       "inlet-outlet": ((FARFIELD, OUTLET), (PERIODIC, PERIODIC), (PERIODIC, PERIODIC)),
       "far field": ((FARFIELD, FARFIELD),) * 3,
       "outlets": ((OUTLET, OUTLET),) * 3}


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("bc", list(BCS), ids=list(BCS))
def test_every_index_stays_inside_its_allocation(lib, shape, bc):
    rc, (max_pop, elements, max_link, table, sources) = check(lib, shape, BCS[bc])
    assert rc == 0
    assert max_pop < elements and max_link < table
    n = shape[0] * shape[1] * shape[2]
    assert elements == 19 * n
    # the walk reaches the last element of the lattice: the checks see the
    # layout's extreme, not a corner of it
    assert max_pop == elements - 1
    assert sources > 0
    for elem_bytes in (4, 2):          # FP32, FP16 / BF16 offsets
        assert max_pop * elem_bytes + elem_bytes <= elements * elem_bytes


def test_largest_32_bit_grid(lib):
    # 4096 x 1024 x 1023 = 4,290,772,992 cells < 2^32: the population index
    # (up to 19 N) needs 64 bits, and the kernels form it in 64 bits
    rc, (max_pop, elements, _, _, _) = check(lib, (4096, 1024, 1023), BCS["periodic"])
    assert rc == 0 and max_pop == elements - 1 and elements > (1 << 32)


def test_refuses_a_grid_beyond_32_bit_cells(lib):
    rc, _ = check(lib, (4096, 1024, 1024), BCS["periodic"])
    assert rc == 5
