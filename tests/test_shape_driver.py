"""Shape-file driver: parsing (CPU) against the reference's own shape files, kept as data fixtures
(tests/golden/tcb.txt and shape.txt are copies of the reference's bin/tcb.txt and bin/shape.txt), and a
GPU run of scaled-down shape lines through the C-ABI, whose classical and Strassen results are both
checked against FP64 sums over the oracle's scatter vectors (tests/sampled.py)."""
import os

import pytest

import shape_driver as sd

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_parse_reference_shape_files():
    with open(os.path.join(GOLDEN, "tcb.txt")) as f:
        tcb = list(sd.parse_shape_lines(f))
    with open(os.path.join(GOLDEN, "shape.txt")) as f:
        shape = list(sd.parse_shape_lines(f))
    assert len(tcb) == 6 and len(shape) == 11
    assert tcb[0] == ("abc", "acd", "db", {"a": 480, "b": 480, "c": 480, "d": 480})
    assert tcb[2] == ("ab", "ac", "cb", {"a": 5136, "b": 5120, "c": 5136})
    assert shape[1] == ("abij", "abef", "efij", dict(a=136, b=136, e=136, f=136, i=112, j=112))
    # m, n, k of a line: the AC, BC and AB groups
    assert sd.group_sizes("acd", "db", "abc", tcb[0][3])[:3] == (480 * 480, 480, 480)
    assert sd.group_sizes("abef", "efij", "abij", shape[0][3])[:3] == (128 * 128, 128 * 128, 128 * 128)


def test_parse_rules():
    assert sd.parse_lengths("a:3;b:40;") == {"a": 3, "b": 40}
    assert sd.parse_lengths("a:3;b:40") == {"a": 3}  # an entry counts only when ';' ends it
    with pytest.raises(sd.ShapeError):
        sd.parse_lengths("ab:3;")
    lines = ["# comment\n", "ab ac cb & a:4;b:5;c:6;\n", "\n", "ab ac cb & a:1;b:1;c:1;\n"]
    assert len(list(sd.parse_shape_lines(lines))) == 1  # the first empty line ends the list
    with pytest.raises(sd.ShapeError):
        list(sd.parse_shape_lines(["ab ac cb & a:4;b:5;\n"]))


@pytest.mark.gpu
def test_driver_runs_scaled_reference_shapes(tmp_path, capsys):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    # the reference's TCCG lines with every length divided by 8 (480 -> 60, 72 -> 9, 5136 -> 642)
    with open(os.path.join(GOLDEN, "tcb.txt")) as f:
        lines = list(sd.parse_shape_lines(f))
    scaled = tmp_path / "small.txt"
    with open(scaled, "w") as f:
        for idx_C, idx_A, idx_B, lengths in lines:
            ls = "".join(f"{c}:{max(1, v // 8)};" for c, v in sorted(lengths.items()))
            f.write(f"{idx_C} {idx_A} {idx_B} & {ls}\n")
    import sampled
    import tblis_strassen as ts

    ts.load()
    for level, impl in ((1, "abc"), (2, "ab"), (1, "naive"), (2, "abc")):
        for idx_C, idx_A, idx_B, lengths in sd.parse_shape_lines(open(scaled)):
            r = sd.run_line(ts, torch, idx_C, idx_A, idx_B, lengths, 1, level, impl, torch.float64, 0, None,
                            keep=True)
            assert r["err2"] <= 1e-24, (idx_C, idx_A, idx_B, level, impl, r)
            # both results against FP64 sums over the oracle's scatter vectors (one call: C = A.B)
            case = dict(idx_A=idx_A, idx_B=idx_B, idx_C=idx_C, lengths=lengths,
                        **{f"shape_{x}": [lengths[c] for c in ix] for x, ix in zip("ABC", (idx_A, idx_B, idx_C))})
            buf = r["buffers"]
            off, ref = sampled.sample(case, buf["A"].cpu().numpy(), buf["B"].cpu().numpy(), nsamp=1024, seed=level)
            idx = torch.from_numpy(off).to("cuda")
            for name in ("C", "Cref"):
                err = sampled.rel_err(buf[name][idx].cpu().numpy(), ref)
                assert err <= 1e-13, (idx_C, idx_A, idx_B, level, impl, name, err)
    sd.main(["-file", str(scaled), "-niter", "1", "-level", "1", "-impl", "1", "-notrash"])
    rows = [r for r in capsys.readouterr().out.splitlines() if r.strip()]
    assert len(rows) == len(lines) and all(len(r.split()) == 9 for r in rows)
