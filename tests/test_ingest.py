"""Model file ingest on the CPU: sx_convert_hmm_file parses and configures an HMMER3 file
(text or gzip) on host threads and writes the packed profile cache.  The cache is read
back here by its documented layout and its tables compared bit for bit with the oracle's
configure (model.rs:146-245) of the oracle's own parse (parse.rs); the thread count and
compression must not change a byte; a malformed file fails with the one-pass reader's
message (parse.rs:104, `<file> line <n>: ...`)."""
import gzip
import os
import struct

import numpy as np
import pytest

from test_search_host import GOLDEN


def read_cache(path):
    """The packed profile cache (sx_ingest.h): header, then one record per profile."""
    b = open(path, "rb").read()
    assert b[:8] == b"SXPCACHE"
    ver, zero, count = struct.unpack_from("<IIQ", b, 8)
    assert ver == 1 and zero == 0
    off, out = 24, []
    for _ in range(count):
        (sz,) = struct.unpack_from("<I", b, off)
        off += 4
        rec = b[off:off + sz]
        off += sz
        m, tbm, ln0, ln1, ln2 = struct.unpack_from("<ifIII", rec, 0)
        flags = rec[20]
        ga, tc, nc = (struct.unpack_from("<2f", rec, 24 + 8 * i) for i in range(3))
        stats = struct.unpack_from("<6d", rec, 48)
        compo = np.frombuffer(rec, np.float32, 20, 96)
        p = 176
        strs = []
        for ln in (ln0, ln1, ln2):
            if ln == 0xFFFFFFFF:
                strs.append(None)
            else:
                strs.append(rec[p:p + ln].decode())
                p += ln
        p = (p + 3) // 4 * 4
        w = m + 1
        tabs = {}
        for k, n in (("msc", 25 * w), ("tsc", 8 * w), ("mo", 25 * w), ("tp", 8 * w)):
            tabs[k] = np.frombuffer(rec, np.float32, n, p)
            p += 4 * n
        assert p == sz
        out.append(dict(m=m, msv_tbm=tbm, name=strs[0], acc=strs[1], desc=strs[2], flags=flags, ga=ga, tc=tc,
                        nc=nc, stats=stats, compo=compo, **tabs))
    assert off == len(b)
    return out


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("fn", ["tiny.hmm", "hmmer34_pair.hmm"])
def test_cache_holds_the_oracles_configured_tables(orc, sxa, tmp_path, fn):
    src = os.path.join(GOLDEN, fn)
    dst = str(tmp_path / "p.sxpc")
    n = sxa.convert_hmm_file(src, dst, threads=4)
    hmms = orc.parse_hmm_file(src)
    recs = read_cache(dst)
    assert n == len(hmms) == len(recs) > 0
    for h, r in zip(hmms, recs):
        p = orc.configure(h)
        assert (r["name"], r["acc"], r["desc"], r["m"]) == (h.name, h.acc, h.desc, h.m)
        for k in ("msc", "tsc", "mo", "tp"):
            assert np.array_equal(bits(r[k]), bits(getattr(p, k).ravel())), k
        assert bits(r["msv_tbm"]) == bits(p.msv_tbm)
        assert np.array_equal(bits(r["compo"]), bits(p.compo))
        for i, (key, c) in enumerate((("GA", h.ga), ("TC", h.tc), ("NC", h.nc))):
            assert bool(r["flags"] & (1 << i)) == (c is not None), key
            if c is not None:
                assert r[key.lower()] == tuple(np.float32(v) for v in c)
        for i, key in enumerate(("MSV", "VITERBI", "FORWARD")):
            assert bool(r["flags"] & (8 << i)) == (key in h.stats)
            if key in h.stats:
                assert r["stats"][2 * i:2 * i + 2] == tuple(h.stats[key])


@pytest.fixture(scope="module")
def synth_file(tmp_path_factory, sxa):
    from sentinel_x_amd import synth
    rng = np.random.default_rng(17)
    hmms = [synth.make_profile(rng, int(m), f"s{i}") for i, m in enumerate(rng.integers(5, 300, 120))]
    for i, h in enumerate(hmms):
        h.acc = f"PF{i:05d}.1" if i % 3 else None
        h.desc = f"synthetic model {i}" if i % 2 else None
        h.stats = {"MSV": (-8.0 - i / 100, 0.7), "VITERBI": (-9.0, 0.7), "FORWARD": (-3.0, 0.7)}
    d = tmp_path_factory.mktemp("ingest")
    path = str(d / "s.hmm")
    synth.write_hmm(path, hmms)
    return path, hmms


def test_threads_and_compression_do_not_change_the_cache(sxa, synth_file, tmp_path):
    path, hmms = synth_file
    gz = str(tmp_path / "s.hmm.gz")
    with open(path, "rb") as f, gzip.open(gz, "wb") as g:
        g.write(f.read())
    outs = []
    for src, th in ((path, 1), (path, 8), (gz, 3)):
        dst = str(tmp_path / f"c{len(outs)}.sxpc")
        assert sxa.convert_hmm_file(src, dst, threads=th) == len(hmms)
        outs.append(open(dst, "rb").read())
    assert outs[0] == outs[1] == outs[2]
    recs = read_cache(str(tmp_path / "c0.sxpc"))
    assert [r["name"] for r in recs] == [h.name for h in hmms]
    assert [r["acc"] for r in recs] == [h.acc for h in hmms]
    assert [r["m"] for r in recs] == [h.m for h in hmms]
    # a cache read back and written again is the same file
    again = str(tmp_path / "again.sxpc")
    assert sxa.convert_hmm_file(str(tmp_path / "c0.sxpc"), again) == len(hmms)
    assert open(again, "rb").read() == outs[0]


def test_header_slashes_and_trailing_blanks_read_as_one_pass(sxa, synth_file, tmp_path):
    """A '//' line inside a header is an unknown tag to the one-pass reader (parse.rs:
    header tags other than the known ones are skipped), and blank lines after the last
    model end the file; the threaded reader must agree on both."""
    path, hmms = synth_file
    lines = open(path).read().split("\n")
    names = [i for i, l in enumerate(lines) if l.startswith("NAME  ")]
    lines.insert(names[3] + 1, "//")
    odd = str(tmp_path / "odd.hmm")
    open(odd, "w").write("\n".join(lines) + "\n\n   \n")
    ref = str(tmp_path / "ref.sxpc")
    sxa.convert_hmm_file(path, ref)
    got = str(tmp_path / "odd.sxpc")
    assert sxa.convert_hmm_file(odd, got, threads=6) == len(hmms)
    assert open(got, "rb").read() == open(ref, "rb").read()


def test_malformed_files_fail_with_the_line_of_the_one_pass_reader(sxa, synth_file, tmp_path):
    path, hmms = synth_file
    lines = open(path).read().split("\n")
    # record 40's node 3: a wrong node index
    starts = [i for i, l in enumerate(lines) if l.startswith("HMMER3/f")]
    r = starts[40]
    node = next(i for i in range(r, len(lines)) if lines[i].split()[:1] == ["3"])
    lines[node] = lines[node].replace("      3   ", "      4   ", 1)
    bad = str(tmp_path / "bad.hmm")
    open(bad, "w").write("\n".join(lines))
    with pytest.raises(sxa.SxError) as e:
        sxa.convert_hmm_file(bad, str(tmp_path / "bad.sxpc"), threads=8)
    assert f"{bad} line {node + 1}: node index 4, expected 3" in str(e.value)
    assert not os.path.exists(str(tmp_path / "bad.sxpc"))
    # a good file cut inside its last model
    good = open(path).read().split("\n")
    cut = str(tmp_path / "cut.hmm")
    open(cut, "w").write("\n".join(good[:starts[-1] + 30]) + "\n")
    with pytest.raises(sxa.SxError) as e:
        sxa.convert_hmm_file(cut, str(tmp_path / "cut.sxpc"), threads=8)
    assert f"{cut}: unexpected end of file (" in str(e.value)


def test_damaged_gzip_is_an_error_not_an_end_of_file(sxa, synth_file, tmp_path):
    """A gzip stream that stops decompressing (here: every model decompresses, then the
    trailer's CRC does not match) is a read error, as flate2's is to the reference
    (parse.rs:48-49), not a shorter file."""
    path, hmms = synth_file
    data = gzip.compress(open(path, "rb").read())
    bad = bytearray(data)
    bad[-8] ^= 0xFF  # the CRC-32 of the trailer
    gz = str(tmp_path / "crc.hmm.gz")
    open(gz, "wb").write(bytes(bad))
    for th in (1, 8):
        with pytest.raises(sxa.SxError, match="reading"):
            sxa.convert_hmm_file(gz, str(tmp_path / f"crc{th}.sxpc"), threads=th)
        assert not os.path.exists(str(tmp_path / f"crc{th}.sxpc"))
    # the same stream intact reads in full
    good = str(tmp_path / "ok.hmm.gz")
    open(good, "wb").write(data)
    assert sxa.convert_hmm_file(good, str(tmp_path / "ok.sxpc"), threads=8) == len(hmms)


def test_empty_file_and_damaged_cache(sxa, synth_file, tmp_path):
    empty = str(tmp_path / "empty.hmm")
    open(empty, "w").write("\n\n")
    assert sxa.convert_hmm_file(empty, str(tmp_path / "e.sxpc")) == 0
    assert read_cache(str(tmp_path / "e.sxpc")) == []
    path, hmms = synth_file
    good = str(tmp_path / "g.sxpc")
    sxa.convert_hmm_file(path, good)
    data = open(good, "rb").read()
    short = str(tmp_path / "short.sxpc")
    open(short, "wb").write(data[:len(data) // 2])
    with pytest.raises(sxa.SxError, match="truncated or damaged profile cache"):
        sxa.convert_hmm_file(short, str(tmp_path / "x.sxpc"))
