"""Device random streams vs the CPU oracle: bit-exact.

xorshift128 (kernel.c:214-239) and RANLUXCL (pyopencl-ranluxcl.cl:363-725,
893-920) as compiled into libfeynsimul.so for gfx950.
"""
import numpy as np
import pytest

from oracle import oracle as O
from feynsimul_amd.physical_systems import HarmOsc

pytestmark = pytest.mark.gpu


def _case(msg):
    print("[case] " + msg, flush=True)


@pytest.fixture(scope="module")
def olib():
    return O.build(HarmOsc(), 64, 5.0, operators=("x1*x1",))


@pytest.fixture(scope="module")
def olib64():
    return O.build(HarmOsc(), 64, 5.0, operators=("x1*x1",), double=True)


def test_seed_rows_match_splitmix(gpu):
    import torch
    from feynsimul_amd import _native
    _case("fs_seed_xorshift 1000 rows")
    n, seed, first = 1000, 123456789, 77
    rows = torch.zeros((n, 4), dtype=torch.int32, device=gpu)
    _native.check(_native.lib().fs_seed_xorshift(rows.data_ptr(), n, seed, first, None))
    torch.cuda.synchronize()
    got = rows.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, O.xorshift_rows(seed, first, n))


@pytest.mark.parametrize("ndraw", [1, 257])
def test_xorshift_stream_bit_exact(gpu, olib, olib64, ndraw):
    import torch
    from feynsimul_amd import _native
    _case("xorshift draws %d" % ndraw)
    rows = O.xorshift_rows(42, 0, 512)
    rows[0] = [0, 0, 0, 1]          # small states exercise the randFloat edge cases
    rows[1] = [0xFFFFFFFF] * 4
    d_rows = torch.from_numpy(rows.view(np.int32)).to(gpu)
    n = rows.shape[0]
    oi = torch.zeros((n, ndraw), dtype=torch.int32, device=gpu)
    of = torch.zeros((n, ndraw), dtype=torch.float32, device=gpu)
    od = torch.zeros((n, ndraw), dtype=torch.float64, device=gpu)
    _native.check(_native.lib().fs_xorshift_draw(d_rows.data_ptr(), oi.data_ptr(), of.data_ptr(), od.data_ptr(),
                                                 n, ndraw, None))
    torch.cuda.synchronize()
    ref_i, ref_f = O.xs_stream(olib, rows, ndraw)
    _, ref_d = O.xs_stream(olib64, rows, ndraw)
    np.testing.assert_array_equal(oi.cpu().numpy().view(np.uint32), ref_i)
    np.testing.assert_array_equal(of.cpu().numpy(), ref_f)
    np.testing.assert_array_equal(od.cpu().numpy(), ref_d)


@pytest.mark.parametrize("lux", [0, 1, 2, 3, 4])
def test_ranlux_stream_bit_exact(gpu, olib, lux):
    import torch
    from feynsimul_amd import _native
    _case("ranlux lux %d" % lux)
    seeds = O.ranlux_seeds(99, 0, 256)
    seeds[0] = 0  # all-zero expansion: carry start and the low-bit fill
    d_seeds = torch.from_numpy(seeds.view(np.int64)).to(gpu)
    tab = torch.zeros((256, 28), dtype=torch.int32, device=gpu)
    L = _native.lib()
    _native.check(L.fs_ranlux_init(tab.data_ptr(), d_seeds.data_ptr(), 256, None))
    torch.cuda.synchronize()
    ref_st = O.ranlux_init(olib, seeds)
    np.testing.assert_array_equal(tab.cpu().numpy().view(np.uint8).reshape(256, 112),
                                  ref_st.view(np.uint8).reshape(256, 112))
    nquad = 61  # crosses several 24-value cycles and luxury skips
    out = torch.zeros((256, nquad * 4), dtype=torch.float32, device=gpu)
    _native.check(L.fs_ranlux_draw(tab.data_ptr(), out.data_ptr(), None, 256, nquad, lux, None))
    torch.cuda.synchronize()
    ref = O.ranlux_stream(olib, ref_st, O.nskip_of_lux(lux), nquad)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    np.testing.assert_array_equal(tab.cpu().numpy().view(np.uint8).reshape(256, 112),
                                  ref_st.view(np.uint8).reshape(256, 112))


def test_ranlux_double_bit_exact(gpu, olib):
    import torch
    from feynsimul_amd import _native
    _case("ranlux64 lux 2")
    seeds = O.ranlux_seeds(5, 0, 64)
    d_seeds = torch.from_numpy(seeds.view(np.int64)).to(gpu)
    tab = torch.zeros((64, 28), dtype=torch.int32, device=gpu)
    L = _native.lib()
    _native.check(L.fs_ranlux_init(tab.data_ptr(), d_seeds.data_ptr(), 64, None))
    out = torch.zeros((64, 40 * 4), dtype=torch.float64, device=gpu)
    _native.check(L.fs_ranlux_draw(tab.data_ptr(), None, out.data_ptr(), 64, 40, 2, None))
    torch.cuda.synchronize()
    st = O.ranlux_init(olib, seeds)
    ref = O.ranlux_stream(olib, st, O.nskip_of_lux(2), 40, double=True)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)


def _xor_bits(vals):
    v = np.ascontiguousarray(vals)
    if v.dtype == np.float64:
        b = v.view(np.uint64)
        w = (b & np.uint64(0xFFFFFFFF)).astype(np.uint32) ^ (b >> np.uint64(32)).astype(np.uint32)
    else:
        w = v.view(np.uint32)
    return np.bitwise_xor.reduce(w, axis=1)


@pytest.mark.parametrize("dbl", [False, True])
def test_stream_benchmark_xorshift_checksums(gpu, olib, olib64, dbl):
    """fs_rng_bench (the reference's RNG timing kernel): the XOR of the bits of every stream's
    `count` uniforms and the advanced rows equal the oracle's stream."""
    import torch
    from feynsimul_amd import _native
    _case("stream benchmark xorshift f64=%d" % dbl)
    n, count = 320, 1001
    rows = O.xorshift_rows(8, 0, n)
    d_rows = torch.from_numpy(rows.view(np.int32).copy()).to(gpu)
    check = torch.zeros(n, dtype=torch.int32, device=gpu)
    _native.check(_native.lib().fs_rng_bench(_native.FS_RNG_XORSHIFT, 0, int(dbl), d_rows.data_ptr(),
                                             check.data_ptr(), n, count, None))
    torch.cuda.synchronize()
    ints, flt = O.xs_stream(olib64 if dbl else olib, rows, count)
    np.testing.assert_array_equal(check.cpu().numpy().view(np.uint32), _xor_bits(flt))
    # the rows advanced by `count` steps: their next word is the oracle's draw count + 1
    nxt, _ = O.xs_stream(olib, rows, count + 1)
    got_next, _ = O.xs_stream(olib, d_rows.cpu().numpy().view(np.uint32), 1)
    np.testing.assert_array_equal(got_next[:, 0], nxt[:, count])


@pytest.mark.parametrize("lux,dbl,count", [(0, False, 97), (2, False, 1000), (4, False, 250), (2, True, 130)])
def test_stream_benchmark_ranlux_checksums(gpu, olib, lux, dbl, count):
    """fs_rng_bench with RANLUX: the buffered wrapper draws ceil(count / 4) float4s (values left in the
    buffer at the end are dropped, kernel.c:180-207); checksum of the first `count` values and the
    advanced states equal the oracle's."""
    import torch
    from feynsimul_amd import _native
    _case("stream benchmark ranlux lux %d f64=%d count %d" % (lux, dbl, count))
    n = 192
    seeds = O.ranlux_seeds(31, 0, n)
    st = O.ranlux_init(olib, seeds)
    tab = torch.from_numpy(st.view(np.int32).reshape(n, 28).copy()).to(gpu)
    check = torch.zeros(n, dtype=torch.int32, device=gpu)
    _native.check(_native.lib().fs_rng_bench(_native.FS_RNG_RANLUX, lux, int(dbl), tab.data_ptr(),
                                             check.data_ptr(), n, count, None))
    torch.cuda.synchronize()
    nquad = (count + 3) // 4
    ref = O.ranlux_stream(olib, st, O.nskip_of_lux(lux), nquad, double=dbl)
    np.testing.assert_array_equal(check.cpu().numpy().view(np.uint32), _xor_bits(ref[:, :count]))
    np.testing.assert_array_equal(tab.cpu().numpy().view(np.uint8).reshape(n, 112), st.view(np.uint8).reshape(n, 112))
