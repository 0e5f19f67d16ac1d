"""The EVP file (include/sevp.h, sevp_evp_write / sevp_evp_read) against the
reference's own Fortran statements compiled by flang (tests/fortran/evp_io.f90):

  PREP  WRITE(FNO_EVP) AL,AR,AB,AT,AC,RINV,RINV1,IE       prep_GLOBAL.f90:320
  MAIN  READ(FN_EVP) AL,AR,AB,AT,AC,OAC,RINV,RINV1,IE     input_GLOBAL.f90:442

The library's files must equal flang's byte for byte in both layouts, each
side must read the other's, and the reference's own mismatch (MAIN's READ of
PREP's record fails: the record is one OAC short) must show, while the
library's reader accepts either layout.  Records above 2 GiB are split into
subrecords by gfortran's convention; no compiler here writes those, so the
framing is checked structurally with a small split size.  Host-only calls:
no GPU is used.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from musoc_solver_amd import sevp as S

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = next((c for c in (shutil.which("flang"), "/opt/rocm/llvm/bin/flang")
              if c and os.path.exists(c)), None)

CASES = {  # name: (i2, j2, nb0, form, nbir)
    "bir": (24, 10, 5, S.FORM_BIR, 4),
    "full": (16, 8, 3, S.FORM_FULL_PERIODIC, 0),
    "one-block": (12, 4, 1, S.FORM_FULL, 0),
}


@pytest.fixture(scope="module")
def evp_io(tmp_path_factory):
    if FLANG is None:
        pytest.skip("flang is not installed")
    d = tmp_path_factory.mktemp("evp_io")
    exe = str(d / "evp_io")
    subprocess.run([FLANG, "-O0", os.path.join(HERE, "fortran", "evp_io.f90"), "-o", exe],
                   check=True, cwd=str(d), capture_output=True)
    return exe


def make_record(name, seed):
    i2, j2, nb0, form, nbir = CASES[name]
    rng = np.random.default_rng(seed)
    f32 = lambda: np.asfortranarray(rng.standard_normal((i2, j2)).astype(np.float32))
    coefs = tuple(f32() for _ in range(5))  # AL, AB, AC, AR, AT
    oac = f32()
    s0, s1 = S.evp_shapes(i2, nb0, form, nbir)
    rinv = np.asfortranarray(rng.standard_normal(s0))
    rinv1 = np.asfortranarray(rng.standard_normal(s1))
    ie = np.cumsum(rng.integers(2, 5, nb0)).astype(np.int32)
    return dict(i2=i2, j2=j2, nb0=nb0, form=form, nbir=nbir, coefs=coefs, oac=oac, rinv=rinv,
                rinv1=rinv1, ie=ie, w2=s0[0] * (s0[1] if form != S.FORM_BIR else 1),
                nmat=s0[2] if form == S.FORM_BIR else 1)


def raw_bytes(r):
    al, ab, ac, ar, at = r["coefs"]
    parts = [al, ar, ab, at, ac, r["oac"], r["rinv"], r["rinv1"], r["ie"]]
    return b"".join(np.asarray(p).tobytes(order="F") for p in parts)


def run_io(exe, mode, raw, evp, r):
    args = [exe, mode, str(raw), str(evp)] + [str(v) for v in
                                               (r["i2"], r["j2"], r["nb0"], r["w2"], r["nmat"])]
    return subprocess.run(args, capture_output=True).returncode


def c_write(path, r, oac, max_subrecord=0):
    S.evp_write(str(path), r["coefs"], r["rinv"], r["rinv1"], r["ie"], r["form"], r["nbir"],
                oac=r["oac"] if oac else None, max_subrecord=max_subrecord)


def c_read(path, r, with_oac=None):
    return S.evp_read(str(path), r["i2"], r["j2"], r["nb0"], r["form"], r["nbir"],
                      with_oac=with_oac)


def same(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()


def assert_record_equal(got, r, oac):
    for g, e in zip(got["coefs"], r["coefs"]):
        assert same(g, e)
    assert same(got["rinv"], r["rinv"]) and same(got["rinv1"], r["rinv1"])
    assert same(got["ie"], r["ie"])
    if oac:
        assert same(got["oac"], r["oac"])
    else:
        assert got["oac"] is None


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("oac", [False, True], ids=["prep", "main"])
def test_bytes_equal_flang(evp_io, tmp_path, name, oac):
    r = make_record(name, 1)
    raw = tmp_path / "in.raw"
    raw.write_bytes(raw_bytes(r))
    assert run_io(evp_io, "write-main" if oac else "write-prep", raw, tmp_path / "f.evp", r) == 0
    c_write(tmp_path / "c.evp", r, oac)
    fb, cb = (tmp_path / "f.evp").read_bytes(), (tmp_path / "c.evp").read_bytes()
    assert len(fb) == len(cb) and fb == cb


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("oac", [False, True], ids=["prep", "main"])
def test_each_side_reads_the_other(evp_io, tmp_path, name, oac):
    r = make_record(name, 2)
    raw = tmp_path / "in.raw"
    raw.write_bytes(raw_bytes(r))
    # flang's file through the library's reader
    assert run_io(evp_io, "write-main" if oac else "write-prep", raw, tmp_path / "f.evp", r) == 0
    assert_record_equal(c_read(tmp_path / "f.evp", r), r, oac)
    assert_record_equal(c_read(tmp_path / "f.evp", r, with_oac=oac), r, oac)
    # the library's file through Fortran's READ.  With NB0 = 1 RINV1 has no
    # elements, and flang's READ leaves every item after a zero-size array
    # unset (iostat 0; a flang runtime quirk seen with a three-item probe),
    # so that case is checked only through the library's reader and the
    # byte equality with flang's WRITE.
    if r["rinv1"].size == 0:
        return
    c_write(tmp_path / "c.evp", r, oac)
    out = tmp_path / "out.raw"
    assert run_io(evp_io, "read-main" if oac else "read-prep", out, tmp_path / "c.evp", r) == 0
    back = out.read_bytes()
    want = raw_bytes(r) if oac else raw_bytes(dict(r, oac=np.zeros_like(r["oac"])))
    assert back == want


def test_main_cannot_read_prep_record(evp_io, tmp_path):
    """The reference's pair of statements disagree: MAIN's READ of the record
    PREP writes runs off its end.  The library's reader takes it."""
    r = make_record("bir", 3)
    raw = tmp_path / "in.raw"
    raw.write_bytes(raw_bytes(r))
    assert run_io(evp_io, "write-prep", raw, tmp_path / "prep.evp", r) == 0
    assert run_io(evp_io, "read-main", tmp_path / "out.raw", tmp_path / "prep.evp", r) == 3
    assert_record_equal(c_read(tmp_path / "prep.evp", r), r, False)


def parse_subrecords(buf):
    """(leading, trailing, payload length) for every subrecord of a file."""
    out, at = [], 0
    while at < len(buf):
        (m1,) = struct.unpack(">i", buf[at:at + 4])
        n = abs(m1)
        (m2,) = struct.unpack(">i", buf[at + 4 + n:at + 8 + n])
        out.append((m1, m2, n))
        at += 8 + n
    assert at == len(buf)
    return out


@pytest.mark.parametrize("split", [1000, 4096, 7])
def test_subrecord_framing(tmp_path, split):
    r = make_record("bir", 4)
    p = tmp_path / "split.evp"
    c_write(p, r, True, max_subrecord=split)
    subs = parse_subrecords(p.read_bytes())
    n = sum(s[2] for s in subs)
    assert n == 4 * 6 * r["i2"] * r["j2"] + 8 * (r["rinv"].size + r["rinv1"].size) + 4 * r["nb0"]
    assert len(subs) == -(-n // split) > 1
    for k, (m1, m2, ln) in enumerate(subs):
        last, first = k == len(subs) - 1, k == 0
        assert m1 == (ln if last else -ln)       # continued: negative leading length
        assert m2 == (ln if first else -ln)      # continuation: negative trailing length
        assert ln == (split if not last else n - split * (len(subs) - 1))
    assert_record_equal(c_read(p, r), r, True)


def test_subrecord_exact_multiple(tmp_path):
    r = make_record("one-block", 5)
    p = tmp_path / "a.evp"
    c_write(p, r, False)
    n = len(p.read_bytes()) - 8
    assert n % 4 == 0
    c_write(p, r, False, max_subrecord=n // 4)
    subs = parse_subrecords(p.read_bytes())
    assert [s[2] for s in subs] == [n // 4] * 4
    assert_record_equal(c_read(p, r), r, False)


def test_rejections(tmp_path):
    r = make_record("bir", 6)
    p = tmp_path / "a.evp"
    c_write(p, r, False)
    with pytest.raises(S.SevpError) as e:
        c_read(p, r, with_oac=True)
    assert e.value.code == S.SEVP_E_IO
    with pytest.raises(S.SevpError, match="expected"):
        c_read(p, r, with_oac=True)
    with pytest.raises(S.SevpError, match="matches neither layout"):
        S.evp_read(str(p), r["i2"], r["j2"] + 1, r["nb0"], r["form"], r["nbir"])
    buf = p.read_bytes()
    (tmp_path / "cut.evp").write_bytes(buf[:-9])
    with pytest.raises(S.SevpError, match="well-formed"):
        c_read(tmp_path / "cut.evp", r)
    bad = bytearray(buf)
    bad[-4:] = struct.pack(">i", 12)  # trailing length disagrees with the leading one
    (tmp_path / "bad.evp").write_bytes(bytes(bad))
    with pytest.raises(S.SevpError, match="well-formed"):
        c_read(tmp_path / "bad.evp", r)
    with pytest.raises(S.SevpError, match="longer than"):
        S.evp_read(str(p), r["i2"], r["j2"] - 1, r["nb0"], r["form"], r["nbir"], with_oac=False)
    huge = bytearray(buf)
    huge[:4] = struct.pack(">i", 2**31 - 1)  # a corrupt length is refused, not allocated
    (tmp_path / "huge.evp").write_bytes(bytes(huge))
    with pytest.raises(S.SevpError, match="longer than"):
        c_read(tmp_path / "huge.evp", r)
    with pytest.raises(S.SevpError, match="cannot open"):
        c_read(tmp_path / "missing.evp", r)
    with pytest.raises(S.SevpError, match="NBIR"):
        S.evp_read(str(p), 25, r["j2"], r["nb0"], S.FORM_BIR, 4)


def test_oracle_influence_round_trip(oracle, tmp_path):
    """PREP's output for a real operator (the oracle's BIR strips) through the
    file and back, bitwise."""
    from musoc_solver_amd import synth
    case = synth.make_case(48, 16, 12, metric="latlon", lat_max=70.0)
    ie = synth.reference_ie(16, synth.nb0_for_height(16, 2))
    r_o, r1_o, _ = oracle.bir_prep(case.coefs, ie, 8)
    p = tmp_path / "prep.evp"
    S.evp_write(str(p), case.coefs, r_o, r1_o, ie, S.FORM_BIR, 8)
    got = S.evp_read(str(p), 48, 16, len(ie), S.FORM_BIR, 8)
    assert got["oac"] is None
    assert same(got["rinv"], np.asfortranarray(r_o)) and same(got["rinv1"], np.asfortranarray(r1_o))
    for g, e in zip(got["coefs"], case.coefs):
        assert same(g, np.asfortranarray(e, dtype=np.float32))
    assert np.array_equal(got["ie"], np.asarray(ie, np.int32))


def test_fortran_binding(evp_io, tmp_path):
    """The sevp_hip module's sevp_evp_read / sevp_evp_write from Fortran: a
    record written by PREP's statement is read through the binding (no OAC
    reported) and written back in MAIN's layout, which MAIN's READ accepts."""
    from musoc_solver_amd import build as B
    if not os.path.exists(B.FLIB):
        pytest.skip("libsevp_fortran.so is not built")
    exe = str(tmp_path / "evp_bind")
    subprocess.run([FLANG, "-O0", f"-I{B.LIBDIR}", os.path.join(HERE, "fortran", "evp_bind.f90"),
                    "-o", exe, f"-L{B.LIBDIR}", "-lsevp_fortran", "-lsevp_hip",
                    f"-Wl,-rpath,{B.LIBDIR}"], check=True, cwd=str(tmp_path), capture_output=True)
    for name in ("bir", "full"):
        r = make_record(name, 7)
        r0 = dict(r, oac=np.zeros_like(r["oac"]))
        raw = tmp_path / "in.raw"
        raw.write_bytes(raw_bytes(r))
        assert run_io(evp_io, "write-prep", raw, tmp_path / "prep.evp", r) == 0
        args = [exe, str(tmp_path / "prep.evp"), str(tmp_path / "b.raw"), str(tmp_path / "main.evp")]
        args += [str(v) for v in (r["i2"], r["j2"], r["nb0"], r["form"], r["nbir"])]
        out = subprocess.run(args, capture_output=True, text=True, check=True).stdout
        assert "had_oac=0" in out
        assert (tmp_path / "b.raw").read_bytes() == raw_bytes(r0)
        assert run_io(evp_io, "read-main", tmp_path / "m.raw", tmp_path / "main.evp", r) == 0
        assert (tmp_path / "m.raw").read_bytes() == raw_bytes(r0)


def test_large_record_chunks_and_odd_splits(evp_io, tmp_path):
    """RINV of 1.3 M values (more than one conversion chunk), written whole
    and with subrecords of an odd length that cut values in two: the payload
    is the same bytes, flang's WRITE gives the same whole record, and both
    files read back exactly."""
    i2, j2, nb0 = 256, 40, 20
    CASES["big"] = (i2, j2, nb0, S.FORM_FULL_PERIODIC, 0)
    try:
        r = make_record("big", 8)
    finally:
        del CASES["big"]
    assert r["rinv"].size > (1 << 20)
    whole, split = tmp_path / "whole.evp", tmp_path / "split.evp"
    c_write(whole, r, False)
    c_write(split, r, False, max_subrecord=3_000_001)
    wb, sb = whole.read_bytes(), split.read_bytes()
    subs = parse_subrecords(sb)
    assert len(subs) == -(-(len(wb) - 8) // 3_000_001) > 2
    payload, at = b"", 0
    for _, _, n in subs:
        payload += sb[at + 4:at + 4 + n]
        at += 8 + n
    assert payload == wb[4:-4]
    raw = tmp_path / "in.raw"
    raw.write_bytes(raw_bytes(r))
    assert run_io(evp_io, "write-prep", raw, tmp_path / "f.evp", r) == 0
    assert (tmp_path / "f.evp").read_bytes() == wb
    assert_record_equal(c_read(whole, r), r, False)
    assert_record_equal(c_read(split, r), r, False)
