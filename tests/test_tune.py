"""Tuning-table files on the CPU: parse, install, save, clear and reject
malformed lines (no device calls)."""
import pytest


@pytest.fixture()
def T():
    import crypto_amd.tune as T

    T.clear()
    yield T
    T.clear()


def test_load_save_clear(T, tmp_path):
    p = tmp_path / "a.tune"
    p.write_text("# comment\n\nmsm 2 0 20 16 8\nmsm 1 1 12 9 0  # window only\nmsm 2 2 18 14 4\nntt 0 20 7 7 6\n"
                 "ntt 1 11 1 10 0\nmsm 0 0 16 11 4\n")
    T.load(str(p))
    assert T.msm_cell(2, 0, 20) == (16, 8)
    assert T.msm_cell(1, 1, 12) == (9, 0)
    assert T.window_bits(2, (1 << 18) + 5, T.MODE_GLV) == 14
    assert T.window_bits(1, 5000, T.MODE_BATCHED) == 9
    assert T.window_bits(1, 3000, T.MODE_BATCHED) == 11 - 4  # rule: floor(log2 n) - 4
    assert T.msm_cell(0, 0, 16) == (11, 4) and T.window_bits(0, 70000) == 11  # secp256k1 single MSMs
    assert T.ntt_split(0, 20) == (7, 7, 6)
    assert T.ntt_split(1, 11) == (1, 10, 0)
    # built-in: one pass up to 2^11, then two balanced passes up to 2^20, then three
    assert T.ntt_split(0, 21) == (7, 7, 7)
    assert T.ntt_split(0, 16) == (8, 8, 0)
    assert T.ntt_split(1, 9) == (9, 0, 0) and T.ntt_split(0, 11) == (11, 0, 0)
    assert T.ntt_split(1, 11) == (1, 10, 0)  # the loaded cell
    assert T.ntt_split(1, 23) == (8, 8, 7) and T.ntt_split(0, 26) == (9, 9, 8)
    q = tmp_path / "b.tune"
    T.save(str(q))
    lines = [ln for ln in q.read_text().splitlines() if ln and not ln.startswith("#")]
    assert sorted(lines) == sorted(["msm 2 0 20 16 8", "msm 1 1 12 9 0", "msm 2 2 18 14 4", "ntt 0 20 7 7 6",
                                    "ntt 1 11 1 10 0", "msm 0 0 16 11 4"])
    T.clear()
    assert T.msm_cell(2, 0, 20) == (0, 0) and T.ntt_split(0, 20) == (10, 10, 0)
    T.load(str(q))
    assert T.msm_cell(2, 2, 18) == (14, 4)


def test_glv_off_cell(T, tmp_path):
    """A GLV cell of window bits 1 records that the plain MSM wins at that size
    class: the GLV mode then reports the plain MSM's window bits."""
    p = tmp_path / "g.tune"
    p.write_text("msm 2 0 22 15 16\nmsm 2 2 22 1 0\nmsm 2 2 21 16 8\n")
    T.load(str(p))
    assert T.msm_cell(2, 2, 22) == (1, 0)
    assert T.window_bits(2, 1 << 22, T.MODE_GLV) == 15  # the plain cell
    assert T.window_bits(2, 1 << 21, T.MODE_GLV) == 16
    q = tmp_path / "h.tune"
    T.save(str(q))
    assert "msm 2 2 22 1 0" in q.read_text().splitlines()


def test_secp_glv_cells(T, tmp_path):
    """secp256k1 (curve 0) has single and GLV cells, no batched ones."""
    p = tmp_path / "s.tune"
    p.write_text("msm 0 0 20 15 0\nmsm 0 2 20 14 4\nmsm 0 2 21 1 0\n")
    T.load(str(p))
    assert T.msm_cell(0, 2, 20) == (14, 4)
    assert T.window_bits(0, 1 << 20, T.MODE_GLV) == 14
    assert T.window_bits(0, 1 << 21, T.MODE_GLV) == T.window_bits(0, 1 << 21)  # the GLV-off cell


@pytest.mark.parametrize("line", ["msm 2 0 20 19 8", "msm 2 0 20 4 8", "msm 2 0 20 16 3", "msm 2 0 20 16 32",
                                  "msm 3 0 20 16 8", "msm 2 3 20 16 8", "msm 2 0 32 16 8", "msm 2 0 20 16",
                                  "msm 2 0 20 16 8 1", "ntt 0 20 10 11 0", "ntt 0 20 12 8 0", "ntt 0 20 0 20 0",
                                  "ntt 0 12 4 0 8", "ntt 2 12 6 6 0", "ntt 0 27 9 9 9", "ntx 0 12 6 6 0",
                                  "msm two 0 20 16 8", "msm 2 0 20 1 0", "msm 2 1 20 1 0",
                                  "msm 0 1 16 10 0", "msm 3 2 16 10 0", "msm -1 0 16 10 0"])
def test_malformed_lines_install_nothing(T, tmp_path, line):
    import crypto_amd

    p = tmp_path / "bad.tune"
    p.write_text("msm 1 0 14 10 4\n" + line + "\n")
    with pytest.raises(crypto_amd.KinetError) as e:
        T.load(str(p))
    assert e.value.code == -2
    assert T.msm_cell(1, 0, 14) == (0, 0)  # the valid first line was not installed either


def test_missing_file(T, tmp_path):
    import crypto_amd

    with pytest.raises(crypto_amd.KinetError):
        T.load(str(tmp_path / "absent.tune"))
