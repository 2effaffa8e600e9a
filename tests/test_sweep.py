"""The reference-grid sweep driver (main.cu:236-305 counterpart): grid order
and row format on the CPU, a two-point sweep on the GPU."""
import pytest

import sweep


def test_grid_matches_reference_loops():
    pts = list(sweep.grid())
    assert len(pts) == 17 * 10
    assert pts[0] == (26, 256) and pts[1] == (26, 512) and pts[10] == (52, 256)
    assert pts[-1] == (2080, 131072)


def test_row_format_matches_reference_fprintf():
    # main.cu:296: "%d \t %d \t %f \t %f \t %f\n"
    assert sweep.format_row(26, 256, 0.5, 1.25, 2.0) == \
        "26 \t 256 \t 0.500000 \t 1.250000 \t 2.000000\n"


def test_read_dat_and_compare(tmp_path):
    """Header lines and zero rows (points the reference did not run,
    main.cu:263-271) are skipped; speed-ups are reference t1 / ours."""
    ref = tmp_path / "ref.dat"
    ref.write_text("Computational times\n\nm \t n \t S1 \t S2 \t S3\n"
                   "26 \t 256 \t 0.6 \t 0.7 \t 0.8\n\n2080 \t 131072 \t 0 \t 0 \t 0\n")
    ours = tmp_path / "ours.dat"
    ours.write_text(sweep.format_row(26, 256, 0.1, 0.2, 0.3)
                    + sweep.format_row(2080, 131072, 1.0, 2.0, 3.0))
    assert sweep.read_dat(str(ref)) == {(26, 256): (0.6, 0.7, 0.8)}
    assert set(sweep.read_dat(str(ours))) == {(26, 256), (2080, 131072)}
    table = sweep.compare(str(ours), str(ref), cols=(26, 2080))
    assert "0.600 / 0.100 ms = 6x" in table
    assert "- / 1.000 ms" in table
    assert "at the 1 common points: median 6.0x" in table
    assert "t3 (FP32+refine) over the reference's Splitting 1 at the 1 common points: median 2.0x" in table


@pytest.mark.gpu
def test_sweep_two_points(gpu, tmp_path):
    out = tmp_path / "t.dat"
    print("[case] sweep 26 x {256, 4096}", flush=True)
    assert sweep.main(["--out", str(out), "--runs", "2", "--batches", "26",
                       "--sizes", "256", "4096"]) == 0
    rows = [r.split("\t") for r in out.read_text().splitlines()]
    assert [(int(r[0]), int(r[1])) for r in rows] == [(26, 256), (26, 4096)]
    for r in rows:
        assert len(r) == 5 and all(float(v) > 0 for v in r[2:])
