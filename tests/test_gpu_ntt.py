"""GPU parity: ntt_nr / intt_rn / RSFr on the HIP kernels, bit-exact against
the oracle (oracle/ntt.c) on the same seeded inputs."""
import numpy as np
import pytest

from provekit_amd.inputs import random_fr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def test_forward_ntt_2p16(gpu, oracle):
    # forward NTT of 2^16 uniform Fr, natural order in, bit-reversed out
    v = random_fr(1 << 16, seed=1, stream="ntt-2^16")
    say("ntt_nr 2^16")
    a, b = v.copy(), v.copy()
    gpu.ntt_nr(a, 1 << 16, 1)
    oracle.ntt_nr(b, 1 << 16, 1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("k", list(range(0, 15)) + [18, 20])
def test_ntt_sizes_and_round_trip(gpu, oracle, k):
    n = 1 << k
    v = random_fr(n, seed=100 + k, stream="ntt")
    say(f"ntt_nr 2^{k}")
    a, b = v.copy(), v.copy()
    gpu.ntt_nr(a, n, 1)
    oracle.ntt_nr(b, n, 1)
    assert np.array_equal(a, b)
    gpu.intt_rn(a)  # round_trip_ntt (ntt/src/ntt.rs:415-430)
    assert np.array_equal(a, v)


@pytest.mark.parametrize("rows_log,cols_log", [(0, 3), (3, 2), (10, 1), (12, 3), (6, 6), (16, 2)])
def test_interleaved(gpu, oracle, rows_log, cols_log):
    n = 1 << (rows_log + cols_log)
    v = random_fr(n, seed=rows_log * 100 + cols_log, stream="interleaved")
    say(f"interleaved rows=2^{rows_log} cols=2^{cols_log}")
    a, b = v.copy(), v.copy()
    gpu.ntt_nr(a, 1 << rows_log, 1)
    oracle.ntt_nr(b, 1 << rows_log, 1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("k,groups", [(10, 2), (12, 4), (16, 8), (20, 4), (14, 1 << 14)])
def test_num_groups(gpu, oracle, k, groups):
    n = 1 << k
    v = random_fr(n, seed=k * 7 + groups, stream="groups")
    say(f"ntt_nr 2^{k} num_groups={groups}")
    a, b = v.copy(), v.copy()
    gpu.ntt_nr(a, n, groups)
    oracle.ntt_nr(b, n, groups)
    assert np.array_equal(a, b)


def test_ntt_errors(gpu):
    import provekit_amd as pk
    v = random_fr(12, seed=1)
    with pytest.raises(pk.ProvekitGpuError) as e:
        gpu.ntt_nr(v, 4, 5)
    assert "divisible by num_groups" in str(e.value)
    with pytest.raises(pk.ProvekitGpuError) as e:
        gpu.ntt_nr(v, 12, 1)
    assert e.value.code == pk.ERR_NOT_POW2
    empty = np.zeros((0, 4), np.uint64)
    gpu.ntt_nr(empty, 0, 1)  # ntt_empty (ntt.rs:507-512)


@pytest.mark.parametrize("k", list(range(0, 13)) + [16, 20])
def test_intt_nr(gpu, oracle, k):
    # intt_nr (ntt.rs:333-348): natural-order evaluations -> bit-reversed
    # coefficients; the oracle's restatement is pinned by brute force
    n = 1 << k
    v = random_fr(n, seed=300 + k, stream="intt_nr")
    say(f"intt_nr 2^{k}")
    a, b = v.copy(), v.copy()
    gpu.intt_nr(a)
    oracle.intt_nr(b)
    assert np.array_equal(a, b)
    # and it inverts ntt_nr up to the two bit-reversal permutations
    gpu.reverse_order(a)
    gpu.ntt_nr(a, n, 1)
    gpu.reverse_order(a)
    assert np.array_equal(a, v)


@pytest.mark.parametrize("n,codeword", [(1, 1), (2, 2), (64, 16), (40, 32), (1 << 12, 1 << 12),
                                        (3 << 16, 1 << 17)])
def test_reverse_order(gpu, oracle, n, codeword):
    v = random_fr(n, seed=n + codeword, stream="reverse_order")
    say(f"reverse_order len={n} codeword={codeword}")
    a, b = v.copy(), v.copy()
    gpu.reverse_order(a, codeword)
    oracle.reverse_order(b, codeword)
    assert np.array_equal(a, b)


def test_inverse_device_entry_points(gpu, oracle):
    n = 1 << 14
    v = random_fr(n, seed=77, stream="inverse-device")
    say("intt_nr_device / intt_rn_device / reverse_order_device 2^14")
    for dev_fn, ref_fn in ((gpu.intt_nr_device, oracle.intt_nr), (gpu.intt_rn_device, oracle.intt_rn)):
        d = gpu.DeviceBuffer.from_array(v)
        dev_fn(d, n)
        gpu.synchronize()
        ref = v.copy()
        ref_fn(ref)
        assert np.array_equal(d.download(shape=(n, 4)), ref)
    d = gpu.DeviceBuffer.from_array(v)
    gpu.reverse_order_device(d, n, n // 4)
    gpu.synchronize()
    ref = v.copy()
    oracle.reverse_order(ref, n // 4)
    assert np.array_equal(d.download(shape=(n, 4)), ref)


def test_inverse_errors(gpu):
    import provekit_amd as pk
    v = random_fr(12, seed=2)
    with pytest.raises(pk.ProvekitGpuError) as e:
        gpu.intt_nr(v)
    assert e.value.code == pk.ERR_NOT_POW2
    with pytest.raises(pk.ProvekitGpuError) as e:
        gpu.reverse_order(v, 0)  # 0.is_power_of_two() is false: the reference asserts
    assert e.value.code == pk.ERR_NOT_POW2
    with pytest.raises(pk.ProvekitGpuError) as e:
        gpu.reverse_order(v, 16)  # past the slice
    assert e.value.code == pk.ERR_SIZE
    one = random_fr(1, seed=3)
    w = one.copy()
    gpu.intt_nr(w)  # 1/1 scaling: unchanged
    assert np.array_equal(w, one)


@pytest.mark.parametrize("log_msg,log_extra,M,log_mask", [
    (0, 0, 1, 0), (1, 1, 1, 0), (2, 3, 3, 1), (4, 2, 4, 3), (3, 0, 2, 2), (4, 3, 1, 3), (2, 1, 4, 0),
    (10, 2, 1, 0), (12, 1, 2, 4)])
def test_rs_interleaved_encode(gpu, oracle, log_msg, log_extra, M, log_mask):
    mlen, mask_len = 1 << log_msg, 1 << log_mask
    codeword = (1 << (mlen + mask_len - 1).bit_length()) << log_extra
    msgs = [random_fr(mlen, seed=r, stream=f"rs{log_msg}{log_extra}{M}{log_mask}") for r in range(M)]
    masks = random_fr(M * mask_len, seed=99, stream="rsmask")
    say(f"interleaved_encode msg=2^{log_msg} M={M} mask={mask_len} codeword={codeword}")
    got = gpu.RSFr().interleaved_encode(msgs, masks, codeword)
    exp = oracle.rs_interleaved_encode(msgs, masks, codeword)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("mlen,M,nmask,codeword,outcome", [
    (1922, 1, 1, 4, "panic"), (3, 2, 3, 4, "panic"), (3, 2, 2, 4, "ok"), (2, 2, 1, 4, "ok"),
    (0, 3, 0, 0, "empty"), (1, 1, 0, 0, "panic"), (1, 2, 0, 1, "ok"), (3, 1, 0, 12, "panic")])
def test_rs_interleaved_encode_bounds(gpu, oracle, mlen, M, nmask, codeword, outcome):
    # the reference's own bounds (provekit/common/src/ntt.rs:70-80): panics
    # come back as errors, the rest equal the oracle (pinned on the CPU)
    msgs = [random_fr(mlen, seed=r, stream="rs-bounds") if mlen else np.zeros((0, 4), np.uint64)
            for r in range(M)]
    masks = random_fr(nmask, seed=5, stream="rs-bounds-mask") if nmask else np.zeros((0, 4), np.uint64)
    say(f"interleaved_encode bounds msg={mlen} M={M} masks={nmask} codeword={codeword}")
    if outcome == "panic":
        with pytest.raises(gpu.ProvekitGpuError):
            gpu.RSFr().interleaved_encode(msgs, masks, codeword)
        return
    got = gpu.RSFr().interleaved_encode(msgs, masks, codeword)
    assert np.array_equal(got, oracle.rs_interleaved_encode(msgs, masks, codeword))
    if outcome == "empty":
        assert got.shape == (0, 4)


def test_rs_no_masks_blowup(gpu, oracle):
    # the commit shape: one message, no masks, blowup 4 (num_cosets = 4)
    m = random_fr(1 << 14, seed=3, stream="rs-commit")
    got = gpu.RSFr().interleaved_encode([m], None, 1 << 16)
    exp = oracle.rs_interleaved_encode([m], None, 1 << 16)
    assert np.array_equal(got, exp)


def test_rs_points_generator_next_order(gpu, oracle):
    rs = gpu.RSFr()
    for cw in (1, 2, 64, 1 << 20):
        idx = [0, 1, 2, 3, cw - 1, cw, 5 * cw + 3] if cw > 1 else [0, 1, 7]
        assert np.array_equal(rs.evaluation_points(0, cw, idx), oracle.rs_evaluation_points(cw, idx))
        assert np.array_equal(rs.generator(cw), oracle.root_of_unity(cw))
    assert rs.next_order(1000) == 1024 and rs.next_order((1 << 28) + 1) is None


def _bitrev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2)


@pytest.mark.parametrize("k", [26, 28])
def test_ntt_max_orders_closed_form_and_round_trip(gpu, k):
    """Orders up to the largest RSFr::next_order accepts (2^28): one nonzero
    coefficient c at position j transforms to out[i] = c * w^(j * bitrev(i))
    (the closed form of ntt_nr, checked at sampled positions), and intt_rn
    brings back exactly the delta.  2^26 also round-trips a dense vector."""
    from tests.util import P, to_int
    n = 1 << k
    w = pow(pow(5, (P - 1) >> 28, P), (1 << 28) // n, P)
    rng = np.random.default_rng(k)
    j = int(rng.integers(1, n))
    c = 0x1234567890ABCDEF1122334455667788 % P  # stored limbs, canonical
    a = np.zeros((n, 4), dtype=np.uint64)
    a[j] = [(c >> (64 * q)) & 0xFFFFFFFFFFFFFFFF for q in range(4)]
    say(f"ntt_nr 2^{k} delta at {j}")
    gpu.ntt_nr(a, n, 1)
    picks = [0, 1, 2, n // 2, n - 1] + [int(x) for x in rng.integers(0, n, 48)]
    for i in picks:
        assert to_int(a[i]) == c * pow(w, j * _bitrev(i, k), P) % P, i
    say(f"intt_rn 2^{k}")
    gpu.intt_rn(a)
    nz = np.flatnonzero(a.any(axis=1))
    assert nz.tolist() == [j] and to_int(a[j]) == c
    del a
    if k <= 26:
        v = rng.integers(0, 1 << 63, size=(n, 4), dtype=np.uint64)
        v[:, 3] &= np.uint64((1 << 61) - 1)  # < 2^253 < p
        b = v.copy()
        say(f"ntt_nr + intt_rn 2^{k} dense")
        gpu.ntt_nr(b, n, 1)
        gpu.intt_rn(b)
        assert np.array_equal(b, v)
