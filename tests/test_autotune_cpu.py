"""CPU checks of the autotuner's candidate sets (no GPU)."""

from cutcell import autotune


def test_spmm_candidates_respect_the_wave():
    for vb in (8, 4):
        for k in (1, 2, 4, 8, 16):
            cands = list(autotune.spmm_candidates(k, 300.0, vb))
            assert cands
            for c in cands:
                assert (k // c["cpl"]) * c["nz"] <= 64 and c["cpl"] * vb <= 16
                assert c["grid"] in (256, 512, 768, 1024)
                assert c["unroll"] in (1, 4, 8) or (c["unroll"] == autotune.WINDOW_UNROLL and k == 1)
            # the nonzero-window kernel is a candidate exactly for one right-hand side
            assert any(c["unroll"] == autotune.WINDOW_UNROLL for c in cands) == (k == 1)


def test_batched_candidates_fit_lds():
    for n_max in (100, 2809, 4096, 8192):
        cands = list(autotune.batched_candidates(n_max))
        assert cands
        for c in cands:
            rpt = -(-n_max // c["block_threads"])
            assert rpt <= 16 and 2 * rpt * c["block_threads"] * 8 <= 160 * 1024
