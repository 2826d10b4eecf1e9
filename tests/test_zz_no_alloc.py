"""The allocation-free SpMV entry point (pkgpu_sparse_mul_device_noalloc,
no counterpart in the reference) against pkgpu_sparse_mul_device, which
allocates its scratch: two calls in a row, the second on a smaller matrix
(the fixed scratch keeps the first call's contents), both bit-exact, plus
the oracle on the same inputs and the size limit."""
import numpy as np
import pytest

from provekit_amd.inputs import random_fr


def csr_with_long_rows(rng, rows, cols, long_rows, seed):
    """1-4 entries per row at uniform columns, and `long_rows` {row: length}
    rows that take the chunked path (more than 32 entries, some spanning
    several 2048-entry chunks)."""
    starts, cidx = [], []
    for r in range(rows):
        starts.append(len(cidx))
        k = long_rows.get(r, int(rng.integers(1, 5)))
        cidx.extend(sorted(rng.choice(cols, size=min(k, cols), replace=False).tolist()))
    vals = random_fr(len(cidx), seed=seed, stream="noalloc-csr")
    return np.array(starts, np.uint32), np.array(cidx, np.uint32), vals


def run(gpu, fn, rows, cols, rs, cs, vs, z):
    bufs = [gpu.DeviceBuffer.from_array(x) for x in (rs, cs, vs, z)]
    out = gpu.DeviceBuffer(32 * rows)
    gpu._call(fn, rows, cols, bufs[0].ptr, cs.shape[0], bufs[1].ptr, bufs[2].ptr, bufs[3].ptr,
              out.ptr, None)
    gpu.synchronize()
    return out.download(shape=(rows, 4))


@pytest.mark.gpu
def test_noalloc_spmv_matches_allocating_entry_point(gpu, oracle):
    rng = np.random.default_rng(1234)
    cases = [
        # first call: 2^16 rows, long rows inside one chunk and across several
        (1 << 16, 1 << 16, {7: 40, 1000: 5000, 30000: 2100, 65535: 33}, 21),
        # second call, smaller: the scratch still holds the first call's
        # flags and partials
        (1000, 4096, {3: 3000, 999: 70}, 22),
    ]
    for rows, cols, long_rows, seed in cases:
        rs, cs, vs = csr_with_long_rows(rng, rows, cols, long_rows, seed)
        z = random_fr(cols, seed=seed, stream="noalloc-z")
        print(f"[case] noalloc spmv rows={rows} cols={cols} nnz={cs.shape[0]}", flush=True)
        got = run(gpu, "pkgpu_sparse_mul_device_noalloc", rows, cols, rs, cs, vs, z)
        exp = run(gpu, "pkgpu_sparse_mul_device", rows, cols, rs, cs, vs, z)
        assert np.array_equal(got, exp)
        assert np.array_equal(got, oracle.sparse_mul(rows, cols, rs, cs, vs, z))
    # more entries than the fixed scratch holds: refused on the host, before
    # anything is launched
    small = [gpu.DeviceBuffer(64) for _ in range(5)]
    with pytest.raises(gpu.ProvekitGpuError) as e:
        gpu._call("pkgpu_sparse_mul_device_noalloc", 1, 1, small[0].ptr, (1 << 26) + 1,
                  small[1].ptr, small[2].ptr, small[3].ptr, small[4].ptr, None)
    assert "2^26" in str(e.value)
