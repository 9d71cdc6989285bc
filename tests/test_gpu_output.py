"""The drop-in CLI's printed output against the reference's print
statements restated over the oracle's results (the 100^3/60^3
pair, -g 16 -s 7 -A 30 -N 10).

Every line is compared as text with its numbers cut out, and number by
number: integers, the 3-decimal translations, the printed dmax, grid edge,
centre and origin, and the Overlap fraction and counts must be
byte-identical; FP64 statistics (#MAP SUM line, CC/PCC) within 1e-9
relative; the score statistics of the FP32 search (Std=/Ave=) within 1e-4;
the DOT Score= within 1e-6 (the device sums each voxel's product in FP64, the
reference rounds each voxel's product to float32 first).  The reference's
print statements: VESPER_1_prob_new.py:86-114, :169-172, :639, :660-682,
:745-774, :1381-1392; main.py:155-191.
Also: prob mode with -alpha 0 delegates to the orig-mode search of the two
density maps (main.py:272-281, :17-65)."""
import contextlib
import io
import re

import numpy as np
import pytest

from oracle import vesper_ref as ref
from oracle.build_c import fastvec_c
from vesper_python_amd import synth
from vesper_python_amd import main as cli

pytestmark = pytest.mark.gpu

_NUM = re.compile(r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?")


def _rtol(line):
    if line.startswith("#MAP SUM=") or line.startswith("Overlap="):
        return 1e-9
    if line.startswith("Std="):
        return 1e-4
    if line.startswith("Score="):
        return 1e-6
    return 0.0


def compare_output(got, exp):
    """Line-by-line comparison; returns the number of byte-identical lines."""
    g, e = got.splitlines(), exp.splitlines()
    assert len(g) == len(e), (len(g), len(e), g[-5:], e[-5:])
    same = 0
    for gl, el in zip(g, e):
        if gl == el:
            same += 1
            continue
        assert _NUM.sub("#", gl) == _NUM.sub("#", el), (gl, el)
        tol = _rtol(el)
        for gn, en in zip(_NUM.findall(gl), _NUM.findall(el)):
            if gn == en:
                continue
            # counts and exactly printed quantities never differ
            assert tol > 0, (gl, el, gn, en)
            assert "." in en or "e" in en.lower(), (gl, el, gn, en)
            if el.startswith("Overlap=") and "/" in gl:
                assert gl.split()[1] == el.split()[1], (gl, el)        # Nm/total
            assert abs(float(gn) - float(en)) <= tol * abs(float(en)), (gl, el, tol)
    return same


def _reference_text(argv, paths, pair, s, g, ang, topn):
    """What the reference prints for `main.py orig` on these maps, its
    print statements restated over oracle results."""
    out = [" ".join(str(a) for a in argv)]
    grids = []
    for tag, m in (("A", pair.a), ("B", pair.b)):
        src, dst = ref.set_vox_size(ref.MapGrid(data=np.asarray(m.data, np.float32), xwidth=np.float32(m.width),
                                                orig=np.asarray(m.origin, np.float32)), m.threshold, s)
        out += ["", "#dmax=" + dst.extra["dmax_printed"], "Nvox= {0}, {0}, {0}".format(dst.xdim),
                "cent= " + ", ".join(str(v) for v in dst.cent), "ori= " + ", ".join(str(v) for v in dst.orig)]
        grids.append((src, dst))
    n = ref.unify_dims(grids[0][1], grids[1][1], s)
    out.append("")
    from oracle import build_c
    lib = build_c.load()
    for src, dst in grids:
        d, v = fastvec_c(lib, src.data, src.xwidth, src.orig, n, s, dst.orig, g)
        st = ref.map_stats(d)
        dst.data, dst.vec = d, v
        dst.ave, dst.std, dst.std_norm_ave = st["ave"], st["std"], st["std_norm_ave"]
        out.append("#MAP SUM={sum} COUNT={cnt} AVE={ave} STD={std} STD_norm={std_norm}".format(
            sum=st["dsum"], cnt=st["Nact"], ave=st["ave"], std=st["std"], std_norm=st["std_norm_ave"]))
    out.append("")
    t, q = grids[0][1], grids[1][1]
    # the product's default ranking: rows near the top-N cutoff and every
    # refinement row decided in FP64
    res = ref.search_map_fft(t, q, TopN=topn, ang=ang, mode="VecProduct", prec="ref32", fp64_recheck=True)
    out += ["###Start Searching###", "", "Std= " + str(res["std"]) + " Ave= " + str(res["ave"]), ""]

    def model(prefix, i, angle, trans):
        nt = ref.convert_trans(t.cent, q.cent, ref.euler2rot(angle), trans, q.xwidth, q.xdim)
        return (prefix + str(i) + " Rotation= (" + str(angle[0]) + " " + str(angle[1]) + " " + str(angle[2])
                + ") Translation= (" + "{:.3f}".format(nt[0]) + " " + "{:.3f}".format(nt[1]) + " "
                + "{:.3f}".format(nt[2]) + ")")
    for i, r in enumerate(res["topn"]):
        out.append(model("M", i, res["angles"][r], res["topn_trans"][i]))
    out += ["", ""]
    for i, r in enumerate(res["refined"]):
        out += ["", model("#", i, r["angle"], r["vec_trans"])]
        sc = ref.get_score(t, r["data"], r["vec"], r["vec_trans"], "ref32")
        out.append("Overlap=" + str(float(sc["Nm"]) / float(sc["total"])) + " " + str(sc["Nm"]) + "/"
                   + str(sc["total"]) + " CC=" + str(sc["cc"]) + " PCC=" + str(sc["pcc"]))
        out.append("Score= " + str(sc["score"]))
    return "\n".join(out) + "\n"


@pytest.fixture(scope="module")
def pair100(tmp_path_factory):
    d = tmp_path_factory.mktemp("out100")
    p = synth.make_config_pair("pair100", 1001)
    return synth.write_pair(p, str(d / "p")), p


def test_cli_orig_output_against_reference_format(gpu, pair100):
    paths, pair = pair100
    argv = [paths["A"], paths["B"], pair.a.threshold, pair.b.threshold, 16.0, 7.0, 30.0, 10, False, "V", False]
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        cli.main(["orig", "-a", paths["A"], "-b", paths["B"], "-t", repr(pair.a.threshold), "-T",
                  repr(pair.b.threshold), "-g", "16", "-s", "7", "-A", "30", "-N", "10"])
    exp = _reference_text(argv, paths, pair, 7.0, 16.0, 30.0, 10)
    same = compare_output(buf.getvalue(), exp)
    total = len(exp.splitlines())
    print("[output] byte-identical lines", same, "of", total, flush=True)
    # everything but the FP statistics (2 #MAP SUM lines, the Std= line,
    # 10 Overlap= lines' CC/PCC digits, 10 Score= lines) is byte-identical
    assert same >= total - 2 - 1 - 2 * 10


def test_prob_alpha_zero_delegates_to_orig(gpu, pair100, tmp_path):
    """-alpha 0: the reference ignores the probability arrays and runs the
    orig-mode search of the two density maps (main.py:272-281 ->
    alpha_is_zero, :17-65), printing what the orig command prints after its
    argument line, except the two blank lines around fastVEC."""
    paths, pair = pair100
    rng = np.random.default_rng(5)
    synth.add_prob_channels(rng, pair.a)
    synth.add_prob_channels(rng, pair.b)
    ppaths = synth.write_pair(pair, str(tmp_path / "pp"))
    common = ["-t", repr(pair.a.threshold), "-T", repr(pair.b.threshold), "-s", "7", "-A", "30", "-N", "4"]
    bo = io.StringIO()
    with contextlib.redirect_stdout(bo):
        ro = cli.main(["orig", "-a", ppaths["A"], "-b", ppaths["B"], *common])
    bp = io.StringIO()
    with contextlib.redirect_stdout(bp):
        rp = cli.main(["prob", "-a", ppaths["A"], "-npa", ppaths["A_prob"], "-b", ppaths["B"], "-npb",
                       ppaths["B_prob"], *common, "-alpha", "0.0"])
    assert [tuple(r["angle"]) for r in rp] == [tuple(r["angle"]) for r in ro]
    assert [r["vec_trans"] for r in rp] == [r["vec_trans"] for r in ro]
    assert [r["vec_score"] for r in rp] == [r["vec_score"] for r in ro]
    assert all("prob_score" not in r for r in rp)
    lo = bo.getvalue().splitlines()[1:]
    lp = bp.getvalue().splitlines()[1:]
    # orig prints a blank line before and after the two fastVEC calls
    # (main.py:183, :186); alpha_is_zero does not (main.py:50-52)
    i = next(k for k, l in enumerate(lo) if l.startswith("#MAP SUM="))
    assert lo[i - 1] == "" and lo[i + 2] == ""
    assert lp == lo[:i - 1] + lo[i:i + 2] + lo[i + 3:]
    assert bp.getvalue().splitlines()[0].endswith(" 0.0 -10 -10 -10 -10")
