"""The reference's log scraper restated: data2csv.sh's Seek() (data2csv.sh:1-48)
reads a solver log line by line with grep / cut / sed.  Each field below
applies the same pipeline to the text, so a test can check that this
build's `Solver` output parses into the fields the reference's CSV expects.

    grep PAT          -> the lines containing PAT (case-sensitive substring)
    head -1 / tail -1 -> first / last of them
    cut -d ',' -f1    -> the text before the first comma
    sed 's/[^0-9.]//g'-> only digits and dots kept
    cut -c 2-         -> drop the first character
"""
import re


def _grep(text, pat):
    return [ln for ln in text.splitlines() if pat in ln]


def _digits(s, keep_dot=True):
    return re.sub(r"[^0-9.]" if keep_dot else r"[^0-9]", "", s)


def _f1(ln):
    return ln.split(",", 1)[0]


def seek(text):
    """The fields data2csv.sh's Seek() prints, as strings (None when the
    grep matches nothing, where the shell would print an empty field)."""
    def first(pat):
        g = _grep(text, pat)
        return g[0] if g else None

    def last(pat):
        g = _grep(text, pat)
        return g[-1] if g else None

    out = {}
    gpu = first("CUDA :")
    mult = first("MultiGPU :")
    out["is_gpu"] = gpu is not None and gpu.split(":", 1)[1].replace(" ", "") == "On"
    out["is_multigpu"] = mult is not None and mult.split(":", 1)[1].replace(" ", "") == "On"
    t = first("time")
    out["Total"] = _digits(t) if t is not None else None
    out["Converge"] = "\n".join(_grep(text, "converge")) or None
    loop = _grep(text, "loop")
    out["Loop"] = "\n".join(_digits(ln, keep_dot=False) for ln in loop) or None
    d, m = first("Dot"), first("MV")
    out["CPU-Dot"] = _digits(_f1(d)) if d is not None else None
    out["CPU-MV"] = _digits(_f1(m)) if m is not None else None
    if out["is_gpu"]:
        d, m = last("Dot"), last("MV")
        out["GPU-Dot"] = _digits(_f1(d)) if d is not None else None
        out["GPU-MV"] = _digits(_f1(m)) if m is not None else None
        h2d, d2h = _grep(text, "H2D time"), _grep(text, "D2H time")
        out["GPU-H2D"] = "\n".join(_digits(_f1(ln))[1:] for ln in h2d) or None
        out["GPU-D2H"] = "\n".join(_digits(_f1(ln))[1:] for ln in d2h) or None
        mem = _grep(text, "Memset")
        out["GPU-Memset"] = "\n".join(_digits(_f1(ln)) for ln in mem) or None
    other = _grep(text, "other")
    out["Other"] = "\n".join(_digits(_f1(ln)) for ln in other) or None
    return out
