"""The specialised-kernel registry (csrc/fast_variants_*.hip) names every
variant; the autotuner caches its picks by name and finds them again among a
stage's candidates (one length, element size, transform kind and line/column
form).  Two variants of one form with the same name would make a cached pick
ambiguous, so every (name, lines) pair must be unique.  The names are
rebuilt here from the registration macros' arguments, as fast_impl.h
stringises them."""
import glob
import os
import re

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "2decomp-fft_amd", "csrc")

MACRO = re.compile(r"\b(D2D_VARIANTS|D2D_DTT|D2D_YPACK|D2D_SPLIT_C2C|D2D_SHUF_C2C|D2D_SPLIT_ALL|"
                   r"D2D_C2C)\(([^()]*)\)")
EXPLICIT = re.compile(r"make_variant<\s*(\w+)\s*,\s*\d+\s*,\s*\d+\s*,\s*\d+\s*,\s*(\w+)\s*,\s*"
                      r"(true|false)[^>]*>\(\s*\"([^\"]+)\"\s*\)", re.S)


def _strip_comments(text):
    return re.sub(r"//[^\n]*", "", text)


def registry_entries():
    """[(file, name, lines)] of every registered variant."""
    out = []
    for path in sorted(glob.glob(os.path.join(CSRC, "fast_variants_*.hip"))):
        text = _strip_comments(open(path).read())
        for m in MACRO.finditer(text):
            kind = m.group(1)
            a = [x.strip() for x in m.group(2).split(",")]
            T, N, TX, NT = a[:4]
            base = "%s n=%s tx=%s nt=%s" % (T, N, TX, NT)
            if kind in ("D2D_VARIANTS", "D2D_DTT", "D2D_C2C", "D2D_SPLIT_C2C", "D2D_SPLIT_ALL"):
                lines = a[4] == "true"
            else:
                lines = False  # the y-pencil and in-wave forms are column tiles
            if kind == "D2D_VARIANTS":
                tags = ["c2c", "r2c", "c2r"]
            elif kind == "D2D_DTT":
                tags = ["dct2", "dct3", "dctx", "dct4"]
            elif kind == "D2D_C2C":
                tags = ["c2c"]
            elif kind == "D2D_SPLIT_C2C":
                tags = ["c2c xs w" + a[6]]
            elif kind == "D2D_SPLIT_ALL":
                tags = [k + " xs w" + a[6] for k in ("c2c", "r2c", "c2r")]
            elif kind == "D2D_SHUF_C2C":
                tags = ["c2c xsh w" + a[5]]
            else:  # D2D_YPACK(T, N, TX, NT, RL, XM, WPE, TAG)
                tag = a[7].strip('"')
                tags = [k + tag for k in ("ypk", "yun", "yup")]
            for t in tags:
                out.append((os.path.basename(path), base + " " + t, lines))
        for m in EXPLICIT.finditer(text):
            out.append((os.path.basename(path), m.group(4), m.group(3) == "true"))
    return out


def test_registry_parsed():
    e = registry_entries()
    # every part file registers something, and the explicit re-radixed
    # tiles are seen
    assert {f for f, _, _ in e} == {os.path.basename(p) for p in
                                     glob.glob(os.path.join(CSRC, "fast_variants_*.hip"))}
    names = {n for _, n, _ in e}
    assert "double n=1024 tx=16 nt=1024 c2c xsh w4" in names
    assert "double n=1600 tx=8 nt=1024 c2c xs w4 r8855" in names
    assert "double n=960 tx=4 nt=256 r2c r8815" in names
    assert len(e) > 200


def test_registry_names_unique():
    seen = {}
    dup = []
    for f, n, lines in registry_entries():
        k = (n, lines)
        if k in seen:
            dup.append((n, lines, seen[k], f))
        seen[k] = f
    assert not dup, dup
