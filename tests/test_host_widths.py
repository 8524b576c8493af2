"""Host logic of the block-solve widths (polar_gpu.fit_width): the widest block whose workspace fits
the budget, split into equal chunks, and, when the block is split, a whole number of 128-byte
lines per chunk (16 FP64 / 32 FP32 columns) as long as that keeps 60% of the width."""
import pytest

torch = pytest.importorskip('torch')

from polar_amd import polar_gpu as pg  # noqa: E402

N, P, NA = 4_000_000, 32, 3_973_097


def _per_col(esz, blocked):
    return esz * ((11 if blocked else 4) * N + P * NA) + 64


@pytest.mark.parametrize('cols_fit,want', [(23, 16), (40, 32), (17, 16), (1000, 128), (12, 12), (8, 8)])
def test_fp64_bank_width(cols_fit, want):
    pc = _per_col(8, False)
    w = pg.fit_width(N, P, NA, 256, torch.float64, budget=cols_fit * pc)
    assert w == want
    assert w * pc <= cols_fit * pc or w == 1


@pytest.mark.parametrize('cols_fit,want', [(50, 32), (128, 128), (300, 256), (20, 20)])
def test_fp32_bank_width(cols_fit, want):
    pc = _per_col(4, True)
    assert pg.fit_width(N, P, NA, 256, torch.float32, budget=cols_fit * pc) == want


def test_single_chunk_keeps_its_width():
    pc = _per_col(8, False)
    assert pg.fit_width(N, P, NA, 100, torch.float64, budget=1000 * pc) == 100
    assert pg.fit_width(N, P, NA, 22, torch.float64, budget=1000 * pc) == 22


def test_line_aligned_chunks_cover_the_block():
    pc = _per_col(8, False)
    for fit in range(9, 129):
        w = pg.fit_width(N, P, NA, 256, torch.float64, budget=fit * pc)
        assert 1 <= w <= fit and w <= 128
        if w < 256 and w > 16 and w % 16:
            # kept unaligned only when whole lines would lose more than 40% of the width
            assert (w // 16) * 16 < 0.6 * w
