"""Variants of the FP64 oracle with one planted mistake, for showing that a
pin catches it: the oracle's C sources are copied to a temporary directory,
one textual replacement is applied, and the copy is compiled with the
oracle's own flags (oracle/Makefile).  CPU only; the committed oracle is
never modified."""
import os
import shutil
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CFLAGS = ["-O1", "-std=c11", "-D_GNU_SOURCE", "-fPIC", "-fno-fast-math", "-ffp-contract=off", "-fopenmp",
          "-w"]
SOURCES = ("oracle_carbon.c", "oracle_npd.c")


def build_mutant(workdir, name, source, old, new):
    """Compile the oracle with `old` replaced by `new` in oracle/`source`
    (every occurrence; `old` must occur).  Returns the library path."""
    d = os.path.join(workdir, name)
    os.makedirs(os.path.join(d, "oracle"), exist_ok=True)
    os.makedirs(os.path.join(d, "include"), exist_ok=True)
    shutil.copy(os.path.join(ROOT, "include", "obme.h"), os.path.join(d, "include", "obme.h"))
    for f in SOURCES + ("oracle.h",):
        text = open(os.path.join(ORACLE, f), encoding="utf-8").read()
        if f == source:
            if old not in text:
                raise ValueError(f"mutation {name}: {old!r} not found in {source}")
            text = text.replace(old, new)
        open(os.path.join(d, "oracle", f), "w", encoding="utf-8").write(text)
    out = os.path.join(d, f"liboracle_{name}.so")
    subprocess.run(["gcc", *CFLAGS, "-shared", "-o", out] + [os.path.join(d, "oracle", f) for f in SOURCES]
                   + ["-lm"], check=True)
    return out


def build_many(mutations, source):
    """{name: library path} for {name: (old, new)}, compiled in parallel."""
    from concurrent.futures import ThreadPoolExecutor
    workdir = tempfile.mkdtemp(prefix="oracle_mutants_")
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = {n: ex.submit(build_mutant, workdir, n, source, o, w) for n, (o, w) in mutations.items()}
        return {n: f.result() for n, f in futs.items()}, workdir
