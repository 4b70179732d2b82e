"""gp_score_tests_dosage_noalloc (device buffers in and out, partial sums in
the library's static device array) against gp_score_tests_dosage (host
arrays, its own device buffers) on the same fit state: the two run the same
kernels in the same order, so the statistics agree to the last bit; two calls
in a row, the second with fewer SNP columns, fewer tests and other genes; an
out-of-range index is reported as an error status."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(N, J, Q, M, seed):
    from gaspacho2_amd import lib, synth
    ds = synth.make_dataset(N, J, Q, M, seed=seed, n_numeric=1)
    T0 = lib.init_ta_with_per(ds.M, ds.Q)
    g = lib.GPContext(ds.N, ds.J, ds.M, ds.Q, ds.nh, ds.Yp, ds.Yi, ds.Yx, ds.Zcov, ds.ls, ds.X0, T0)
    g.init_fit()
    for jtr in range(2):
        g.irls_step(jtr)
    return ds, g


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself
    is linked against (the test's own device buffers)."""

    def __init__(self):
        import ctypes as C
        from gaspacho2_amd import lib
        lib.load()
        self.C = C
        import os
        # the runtime libgaspacho2.so is linked against (its rpath is ROCm's lib directory)
        self.rt = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"),
                         mode=C.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def put(self, a: np.ndarray) -> int:
        C = self.C
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        return p.value

    def get(self, p: int, a: np.ndarray) -> np.ndarray:
        assert self.rt.hipMemcpy(a.ctypes.data, self.C.c_void_p(p), a.nbytes, 2) == 0   # device to host
        return a

    def free(self, *ps):
        for p in ps:
            self.rt.hipFree(self.C.c_void_p(p))


def _device_call(g, D, gene, snp):
    hip = _Hip()
    Df = np.ascontiguousarray(np.asfortranarray(D).ravel(order="F"))
    stat = np.full(len(gene), 7.0)
    ptrs = [hip.put(Df), hip.put(np.ascontiguousarray(gene, np.int32)),
            hip.put(np.ascontiguousarray(snp, np.int32)), hip.put(stat)]
    assert hip.rt.hipDeviceSynchronize() == 0
    try:
        g.score_tests_device(D.shape[1], ptrs[0], len(gene), ptrs[1], ptrs[2], ptrs[3])
        return hip.get(ptrs[3], stat)
    finally:
        hip.free(*ptrs)


def test_no_alloc_score_tests_match_the_allocating_entry_point():
    from gaspacho2_amd import lib
    ds, g = _state(3000, 40, 5, 32, seed=2)
    rng = np.random.default_rng(5)
    # first call: 11 SNP columns, 300 tests over every gene
    D = rng.integers(0, 3, (ds.N, 11)).astype(np.uint8)
    gene = rng.integers(0, ds.J, 300).astype(np.int32)
    snp = rng.integers(0, 11, 300).astype(np.int32)
    ref = g.score_tests(D, gene, snp)
    got = _device_call(g, D, gene, snp)
    assert np.all(np.isfinite(ref))
    assert np.array_equal(got, ref)
    # second call, smaller: 4 columns, 37 tests (not a multiple of the chunk)
    D2 = rng.integers(0, 3, (ds.N, 4)).astype(np.uint8)
    gene2 = rng.integers(0, ds.J, 37).astype(np.int32)
    snp2 = rng.integers(0, 4, 37).astype(np.int32)
    ref2 = g.score_tests(D2, gene2, snp2)
    got2 = _device_call(g, D2, gene2, snp2)
    assert np.array_equal(got2, ref2)
    # an index out of range: the kernel's own bounds check turns it into an error status
    snp3 = snp2.copy()
    snp3[5] = 4
    with pytest.raises(lib.GPError):
        _device_call(g, D2, gene2, snp3)
    g.close()


def test_no_alloc_score_tests_at_m256_after_a_narrower_context():
    """A second context with a wider design (P = 257) on the same GPU reuses
    the static array; still bit-identical to the allocating entry point."""
    ds, g = _state(5000, 24, 10, 256, seed=4)
    rng = np.random.default_rng(6)
    D = rng.integers(0, 3, (ds.N, 6)).astype(np.uint8)
    gene = np.repeat(np.arange(ds.J), 6).astype(np.int32)
    snp = np.tile(np.arange(6), ds.J).astype(np.int32)
    ref = g.score_tests(D, gene, snp)
    got = _device_call(g, D, gene, snp)
    assert np.array_equal(got, ref)
    g.close()
