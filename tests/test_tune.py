"""Autotuner host logic (no GPU): the shapes it offers and how it ranks timed shapes."""
from rbffd_gpu_amd import tune as TU


def test_candidates_offer_staged_shapes_only_when_records_fit():
    F = 16
    fits = (128, 765)  # 765 records x 16 FP32 fields = 48960 B
    shapes = list(TU.candidates(F, TU.KIND_RK4, 4, fits, field_bytes=4, fp32_accum=True))
    staged = [s for s in shapes if s[3] in (4, 5)]
    assert (4, 512, 4, 5) in staged and (4, 512, 4, 4) in staged
    assert all(s[1] == 128 * (F // s[0]) for s in staged)  # the block is the tile
    # FP64 accumulation: no matrix-core (share 5) shapes
    assert not any(s[3] == 5 for s in TU.candidates(F, TU.KIND_RK4, 4, fits, field_bytes=4, fp32_accum=False))
    # records past the LDS of a workgroup: no staged shapes at all
    too_big = (128, TU.LDS_BYTES // (F * 8) + 1)
    assert not any(s[3] in (4, 5) for s in TU.candidates(F, TU.KIND_RK4, 4, too_big, field_bytes=8))
    # every direct-gather shape divides the fields and fits its group in the block
    for fpt, block, batch, share in shapes:
        assert F % fpt == 0 and F // fpt <= block and batch in (4, 8)


def test_confirm_reranks_finalists_by_their_best_time():
    first = [(1.0, "a"), (0.8, "b"), (0.7, "c"), (2.0, "d"), (0.9, "e"), (3.0, "f")]
    retimed = {("a",): 0.95, ("b",): 0.6, ("c",): 0.75, ("e",): 0.85}
    calls = []

    def time_shape(*shape):
        calls.append(shape)
        return retimed[shape]

    out = TU._confirm(first, time_shape, finalists=4, rounds=3)
    # the four fastest of the first pass are timed again, three rounds each
    assert sorted(set(calls)) == [("a",), ("b",), ("c",), ("e",)]
    assert len(calls) == 12
    # ranked by the best of all their samples; the rest keep their first-pass order behind them
    assert [r[1:] for r in out] == [("b",), ("c",), ("e",), ("a",), ("d",), ("f",)]
    assert out[0][0] == 0.6 and out[1][0] == 0.7  # c's first sample (0.7) beats its retimes
    assert [r[0] for r in out] == sorted(r[0] for r in out)
