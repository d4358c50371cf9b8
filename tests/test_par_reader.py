"""HITRAN .par / TIPS readers: the oracle pinned by hand-built records, the
C++ reader (lbl_par_read, host code, no GPU) against the oracle on clean and
damaged files, the writer round trip, and from_files (mirrors the reference's
tests/test_data_loader.py TestFromFiles).  CPU only."""

import logging

import numpy as np
import pytest

from oracle import par_oracle as PO
from pylblrtm_amd import data_loader as dl
from pylblrtm_amd import hitran_manager as hm
from pylblrtm_amd import synth
from pylblrtm_amd.spectral_data import SpectralData

LOGGER = "pylblrtm_amd.hitran_manager"


def record(mol=" 21", nu="  700.123456", sref=" 1.234E-20", gair=".0728", gself="0.095",
           elow="  123.4567", nair=" .75", shift="-.002000", tail_len=93):
    """One record from explicit field strings (column layout of
    hitran_manager.py:663-686: ``mol`` is mol_id (2) + iso_id (1)); the tail
    holds the unused columns."""
    head = mol + nu + sref + " 0.000E+00" + gair + gself + elow + nair + shift
    assert len(head) == 67
    return head + " " * tail_len


def write(tmp_path, rows, name="t.par", end="\n"):
    p = tmp_path / name
    p.write_bytes(end.join(rows).encode("utf-8") + end.encode())
    return p


def same(a, b):
    assert set(a) == set(b)
    for k in a:
        assert a[k].dtype == b[k].dtype, k
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


# ------------------------------------------------------------ oracle pins --
def test_oracle_known_record(tmp_path):
    d, bad = PO.par_to_dict(write(tmp_path, [record()]))
    assert bad == []
    assert d["mol_id"].tolist() == [2] and d["iso_id"].tolist() == [1]
    assert d["nu0"][0] == 700.123456 and d["Sref"][0] == 1.234e-20
    assert d["g_air"][0] == 0.0728 and d["g_self"][0] == 0.095
    assert d["Elow"][0] == 123.4567 and d["n_air"][0] == 0.75 and d["shift"][0] == -0.002


def test_oracle_length_rule_and_partial_lines(tmp_path):
    rows = [record(tail_len=88),                      # 155 characters: read
            record(tail_len=87),                      # 154: skipped
            record(sref=" 1.234X-20"),                # bad Sref: mol, iso, nu0 kept
            record(nair=" .7a")]                      # bad n_air: up to Elow kept
    d, bad = PO.par_to_dict(write(tmp_path, rows))
    assert [len(d[n]) for n in PO.NAMES] == [3, 3, 3, 2, 2, 2, 2, 1, 1]
    assert [b[0] for b in bad] == [3, 4]
    assert "could not convert string to float" in bad[0][1]


def test_oracle_newlines_and_errors(tmp_path):
    rows = [record(), record(nu="  701.000000")]
    ref, _ = PO.par_to_dict(write(tmp_path, rows, "a.par"))
    for end in ("\r\n", "\r"):
        d, _ = PO.par_to_dict(write(tmp_path, rows, "b.par", end))
        same(d, ref)
    with pytest.raises(PO.NoLines):
        PO.par_to_dict(write(tmp_path, ["short"], "c.par"))
    with pytest.raises(FileNotFoundError):
        PO.par_to_dict(tmp_path / "missing.par")


# ------------------------------------------------- C++ reader vs the oracle --
def _lines(n, seed):
    rng = np.random.default_rng(seed)
    parts = [synth.random_lines(rng, n // 3 + 1, 500.0, 2500.0, m) for m in (1, 2, 6)]
    out = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    out["iso_id"] = rng.integers(1, 4, out["nu0"].size).astype(np.int32)
    return out


def check_against_oracle(path, caplog):
    ref, bad = PO.par_to_dict(path)
    caplog.clear()
    with caplog.at_level(logging.WARNING, logger=LOGGER):
        got = hm.par_to_dict(path)
    same(got, ref)
    msgs = [(r.args[0], r.args[2] if isinstance(r.args[2], str) else str(r.args[2]))
            for r in caplog.records if r.name == LOGGER]
    assert msgs == [(n, m) for n, m in bad]


def test_reader_matches_oracle_clean(tmp_path, caplog):
    path = hm.dict_to_par(_lines(20000, 1), tmp_path / "clean.par")
    check_against_oracle(path, caplog)


DAMAGE = list(" x_.eE+-nif1\t\x0b\x0c\x1c0") + ["inf", "nan", "1_0", "__", "+", "Infinity"]


# 14,000 lines (2.3 MB) are read by several threads, whose chunk cuts then
# fall inside damaged regions and between the characters of a \r\n
@pytest.mark.parametrize("seed,n", [(5, 14000), (6, 4000), (7, 4000)])
def test_reader_matches_oracle_damaged(tmp_path, caplog, seed, n):
    rng = np.random.default_rng(seed)
    path = hm.dict_to_par(_lines(n, seed), tmp_path / "clean.par")
    rows = path.read_text().splitlines()
    out = []
    for r in rows:
        u = rng.random()
        if u < 0.15:                                  # cut: around the 155 limit
            r = r[: int(rng.integers(140, 170))]
        elif u < 0.55:                                # damage inside the fields
            for _ in range(int(rng.integers(1, 3))):
                i = int(rng.integers(0, 67))
                ins = DAMAGE[int(rng.integers(len(DAMAGE)))]
                r = r[:i] + ins + r[i + len(ins):] if rng.random() < 0.7 else r[:i] + ins + r[i:]
        out.append(r)
    ends = ["\n", "\r\n", "\r"]
    text = "".join(r + ends[int(rng.integers(3))] for r in out)
    p = tmp_path / "damaged.par"
    p.write_bytes(text.encode("ascii"))
    check_against_oracle(p, caplog)


@pytest.mark.parametrize("end", ["\r\n", "\r"])
def test_reader_threaded_line_ends(tmp_path, caplog, end):
    rows = hm.dict_to_par(_lines(30000, 8), tmp_path / "x.par").read_text().splitlines()
    p = tmp_path / "ends.par"
    p.write_bytes((end.join(rows) + end).encode())
    check_against_oracle(p, caplog)


def test_reader_python_number_forms(tmp_path, caplog):
    rows = [record(nu="  7_00.12345"), record(nu="         inf"), record(nu="        -NaN"),
            record(nu=" +700.123456"), record(sref="     1e-20"), record(gair="1._5 "),
            record(gself=" .5  "), record(elow="   1_2.5__"), record(mol="+21"),
            record(mol="-11"), record(mol=" _1"), record(mol=" 2a"), record(elow="\x1c123.4567\x1d"),
            record(shift="1e999999"), record(shift="-1e-9999")]
    check_against_oracle(write(tmp_path, rows), caplog)


def test_non_ascii_file_takes_the_text_path(tmp_path, caplog):
    rows = [record(), record(nu="  70１.123456"), record(tail_len=92) + "é"]
    p = tmp_path / "u.par"
    p.write_bytes("\n".join(rows).encode("utf-8") + b"\n" + record().encode() + b"\xff\n")
    check_against_oracle(p, caplog)


def test_reader_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        hm.par_to_dict(tmp_path / "missing.par")
    with pytest.raises(hm.HITRANDataError, match="No valid lines"):
        hm.par_to_dict(write(tmp_path, ["short line"]))
    with pytest.raises(hm.HITRANDataError):
        hm.par_to_dict(write(tmp_path, [record(nu="   not a nu ")], "bad.par"))


def test_round_trip_rounds_to_the_format(tmp_path):
    lines = _lines(3000, 3)
    got = hm.par_to_dict(hm.dict_to_par(lines, tmp_path / "rt.par"))
    fmt = {"nu0": "{:.6f}", "Sref": "{:.3E}", "g_air": "{:.4f}", "g_self": "{:.3f}",
           "Elow": "{:.4f}", "n_air": "{:.2f}", "shift": "{:.6f}"}
    for k, f in fmt.items():
        np.testing.assert_array_equal(got[k], [float(f.format(v)) for v in lines[k]], err_msg=k)
    np.testing.assert_array_equal(got["mol_id"], lines["mol_id"])
    np.testing.assert_array_equal(got["iso_id"], lines["iso_id"])


def test_txt_to_tips_matches_oracle(tmp_path):
    p = tmp_path / "tips.txt"
    p.write_text("# T Q\n\n   70.0     12.5\n71.0\t13.25  extra\n72 x\nbad 1\n 73.0  14.0\n\r74.0 15.0\r\n",
                 encoding="utf-8")
    t, q = hm.txt_to_tips(p)
    rt, rq = PO.txt_to_tips(p)
    np.testing.assert_array_equal(t, rt)
    np.testing.assert_array_equal(q, rq)
    # hitran_manager.py:955-956 appends T before converting Q: "72 x" keeps
    # its T and drops its Q ("bad 1" fails on T and keeps neither), so the
    # reference returns arrays of different lengths
    assert t.tolist() == [70.0, 71.0, 72.0, 73.0, 74.0]
    assert q.tolist() == [12.5, 13.25, 14.0, 15.0]
    clean = tmp_path / "clean.txt"
    clean.write_text("70 1.5\n# c\n71 2.5e1\n", encoding="utf-8")
    for f in (hm.txt_to_tips, PO.txt_to_tips):
        ct, cq = f(clean)
        assert ct.tolist() == [70.0, 71.0] and cq.tolist() == [1.5, 25.0]
    empty = tmp_path / "empty.txt"
    empty.write_text("# nothing\n")
    with pytest.raises(hm.HITRANDataError):
        hm.txt_to_tips(empty)
    with pytest.raises(FileNotFoundError):
        hm.txt_to_tips(tmp_path / "none.txt")


# ------------------------------------------------------------- from_files --
def _dummy_tips():
    t = np.arange(70.0, 401.0)
    return t, 0.5 * t


def _make_lines(*isos):
    n = 5
    return {"nu0": np.linspace(700.0, 701.0, n * len(isos)), "Sref": np.full(n * len(isos), 1e-20),
            "g_air": np.full(n * len(isos), 0.07), "g_self": np.full(n * len(isos), 0.09),
            "Elow": np.full(n * len(isos), 100.0), "n_air": np.full(n * len(isos), 0.7),
            "shift": np.zeros(n * len(isos)),
            "mol_id": np.full(n * len(isos), 1, dtype=np.int32),
            "iso_id": np.repeat(np.asarray(isos, dtype=np.int32), n)}


def test_from_files_single_iso(monkeypatch):
    monkeypatch.setattr(hm, "par_to_dict", lambda p: _make_lines(1))
    monkeypatch.setattr(hm, "get_molar_mass", lambda m, i: 18.01)
    sd = dl.from_files("dummy.par", tips={(1, 1): _dummy_tips()})
    assert isinstance(sd, SpectralData) and (1, 1) in sd.tips and sd.n_lines == 5
    assert sd.molecule == "H2O"


def test_from_files_missing_tips(monkeypatch):
    monkeypatch.setattr(hm, "par_to_dict", lambda p: _make_lines(1, 2))
    with pytest.raises(ValueError, match="Missing TIPS"):
        dl.from_files("dummy.par", tips={(1, 1): _dummy_tips()})


def test_from_files_tips_path_arrays_and_masses(tmp_path, monkeypatch):
    t, q = _dummy_tips()
    f = tmp_path / "tips.txt"
    f.write_text("\n".join(f"{a:.1f}  {b:.6f}" for a, b in zip(t, q)), encoding="utf-8")
    monkeypatch.setattr(hm, "par_to_dict", lambda p: _make_lines(1, 2))
    calls = []
    monkeypatch.setattr(hm, "get_molar_mass", lambda m, i: calls.append((m, i)) or 18.0 + i)
    sd = dl.from_files("dummy.par", tips={(1, 1): f, (1, 2): (t, q), (7, 1): (t, q)})
    np.testing.assert_allclose(sd.tips[(1, 1)][0], t)
    np.testing.assert_allclose(sd.tips[(1, 2)][1], q)
    assert (7, 1) not in sd.tips                       # entries for absent isotopologues dropped
    assert sorted(calls) == [("H2O", 1), ("H2O", 2)]
    assert sd.molar_masses[(1, 2)] == 20.0
    sd2 = dl.from_files("dummy.par", tips={(1, 1): (t, q), (1, 2): (t, q)},
                        molar_masses={(1, 1): 18.999, (1, 2): 20.1})
    assert sd2.molar_masses[(1, 1)] == 18.999


def test_from_files_unknown_molecule_and_no_mass_table(monkeypatch):
    lines = _make_lines(1)
    lines["mol_id"][:] = 99
    monkeypatch.setattr(hm, "par_to_dict", lambda p: lines)
    with pytest.raises(ValueError, match="not in the internal catalogue"):
        dl.from_files("dummy.par", tips={(99, 1): _dummy_tips()})
    monkeypatch.setattr(hm, "par_to_dict", lambda p: _make_lines(1))
    with pytest.raises(hm.HITRANDataError, match="molar_masses"):
        dl.from_files("dummy.par", tips={(1, 1): _dummy_tips()})


def test_from_files_end_to_end(tmp_path):
    lines = _lines(300, 4)
    par = hm.dict_to_par(lines, tmp_path / "mix.par")
    t, q = _dummy_tips()
    tips = {k: (t, q) for k in set(zip(lines["mol_id"].tolist(), lines["iso_id"].tolist()))}
    masses = {k: 20.0 for k in tips}
    sd = dl.from_files(par, tips=tips, molar_masses=masses)
    assert sd.n_lines == lines["nu0"].size and sd.molecule is None
    np.testing.assert_array_equal(sd.lines["nu0"], [float(f"{v:.6f}") for v in lines["nu0"]])
