"""bsar_range_compress_device_noalloc against bsar_range_compress_device.

The no-allocation entry point runs the same kernels with its buffers in a
fixed per-GPU scratch array, so its output must be bit-identical to the
allocating entry point's on the same input.  Two calls in a row (the second
smaller, so stale scratch contents from the first would show), in both
precisions and on the LDS, split and global filter paths, plus the
over-capacity rejection, which leaves the caller's columns untouched."""
import numpy as np
import pytest

from test_oracle_golden import bits_equal

pytestmark = pytest.mark.gpu

CAPACITY = -6      # BSAR_ERR_CAPACITY


def columns(S, N, seed, dtype):
    rng = np.random.default_rng(seed)
    sig = (rng.normal(size=(N, S)) + 1j * rng.normal(size=(N, S))).astype(dtype)
    ref = rng.normal(size=S) + 1j * rng.normal(size=S)
    return sig, ref


def run(bsar, ctx, sig, ref, precision, conj, noalloc):
    N, S = sig.shape
    b = bsar.DeviceBuffer.from_array(sig)
    try:
        rep = ctx.range_compress_device(b, S, N, ref, conj=conj, precision=precision, noalloc=noalloc)
        out = b.read((N, S), sig.dtype)
    finally:
        b.close()
    return out, rep


# (S, N) pairs: first call, then a smaller one
CASES = {
    "fp64": [(16384, 6), (4096, 3), (32768, 1), (8, 2)],      # split, LDS, global, register-only
    "mixed": [(32768, 5), (1024, 2), (65536, 1), (16, 3)],    # split, LDS, global, register-only
}


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("conj", [True, False])
def test_noalloc_matches_allocating_entry_point(bsar, ctx, form, conj):
    precision = bsar.FP64 if form == "fp64" else bsar.MIXED
    dtype = np.complex128 if form == "fp64" else np.complex64
    for k, (S, N) in enumerate(CASES[form]):
        print(f"noalloc {form} conj={conj} call {k + 1}: S={S} N={N}", flush=True)
        sig, ref = columns(S, N, 100 * S + N + conj, dtype)
        want, _ = run(bsar, ctx, sig, ref, precision, conj, noalloc=False)
        got, rep = run(bsar, ctx, sig, ref, precision, conj, noalloc=True)
        assert rep["launches"] == 1 and rep["pulses"] == N and rep["seconds"] > 0
        fl = np.float64 if form == "fp64" else np.float32
        assert bits_equal(got.view(fl), want.view(fl)), (form, S, N)


def test_noalloc_rejects_sizes_beyond_the_scratch(bsar, ctx):
    S, N = 65536, 8                         # global path: N * S complex128 of scratch alone is 8 MiB
    print(f"noalloc over capacity: S={S} N={N}", flush=True)
    sig, ref = columns(S, N, 5, np.complex128)
    b = bsar.DeviceBuffer.from_array(sig)
    try:
        with pytest.raises(bsar.BsarError) as e:
            ctx.range_compress_device(b, S, N, ref, precision=bsar.FP64, noalloc=True)
        assert e.value.status == CAPACITY
        assert "scratch" in str(e.value)
        assert bits_equal(b.read((N, S), np.complex128).view(np.float64), sig.view(np.float64))
    finally:
        b.close()
    # a call that fits still runs after the rejection
    sig, ref = columns(4096, 2, 6, np.complex128)
    got, _ = run(bsar, ctx, sig, ref, bsar.FP64, True, noalloc=True)
    want, _ = run(bsar, ctx, sig, ref, bsar.FP64, True, noalloc=False)
    assert bits_equal(got.view(np.float64), want.view(np.float64))
