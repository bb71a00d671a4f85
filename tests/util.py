"""Shared helpers for the tests: built binaries, workload generation, oracle runs."""
import importlib.util
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gcsm-bu_amd")
TOOL = os.path.join(PKG, "bin", "csm_tool")
CSMG = os.path.join(PKG, "bin", "csmg")
ORACLE = os.path.join(ROOT, "oracle", "bin", "oracle")
REF_PLAN = os.path.join(ROOT, "oracle", "_ref", "ref_plan")
GOLDEN = os.path.join(ROOT, "tests", "golden")
HEADER = os.path.join(ROOT, "include", "csm.h")


def ensure_built():
    if not (os.path.exists(TOOL) and os.path.exists(ORACLE) and os.path.exists(os.path.join(PKG, "lib", "libcsm.so"))):
        spec = importlib.util.spec_from_file_location("graft_entry", os.path.join(ROOT, "__graft_entry__.py"))
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        m.build()


def ffi():
    spec = importlib.util.spec_from_file_location("csm_ffi", os.path.join(PKG, "csm_ffi.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def autotune():
    spec = importlib.util.spec_from_file_location("autotune", os.path.join(PKG, "autotune.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def run(cmd, timeout=600):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(map(str, cmd))} -> {r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}")
    return r.stdout


def gen(config, out_dir, seed=1, scale=None):
    cmd = [TOOL, "gen", config, str(out_dir), "--seed", str(seed)]
    if scale is not None:
        cmd += ["--scale", str(scale)]
    run(cmd)
    meta = {}
    for line in open(os.path.join(out_dir, "meta.txt")):
        k, v = line.split()
        meta[k] = v
    queries = [os.path.join(out_dir, f"q{k}.graph") for k in range(int(meta["queries"]))]
    return meta, queries


def oracle_stream(data, update, batch, queries, threads=4, directed=False, extra=()):
    cmd = [ORACLE, "stream", str(data), str(update), str(batch)] + [str(q) for q in queries]
    cmd += ["--threads", str(threads)] + list(extra)
    if directed:
        cmd.append("--directed")
    out = run(cmd, timeout=1200)
    res, times = {}, {}
    for line in out.splitlines():
        t = line.split()
        if len(t) == 12 and t[0] == "batch" and t[2] == "query":
            # (positive, negative, positive hash, negative hash)
            res[(int(t[1]), int(t[3]))] = (int(t[9]), int(t[5]), int(t[11], 16), int(t[7], 16))
        elif len(t) == 6 and t[0] == "batch" and t[2] == "edges":
            times[int(t[1])] = float(t[5])
    return res, times


def oracle_count(data, query, directed=False):
    cmd = [ORACLE, "count", str(data), str(query)] + (["--directed"] if directed else [])
    t = run(cmd).split()
    return int(t[1]), int(t[3], 16)


def write_final_graph(data, update, out, directed=False):
    """Applies the whole update stream to the data graph file (plain Python, small inputs).  Edges
    are keyed by their unordered pair, or by the ordered pair when `directed`."""
    vs, es = [], {}

    def key(a, b):
        return (a, b) if directed else (min(a, b), max(a, b))

    for line in open(data):
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            vs.append(line.strip())
        elif t[0] == "e":
            a, b, l = map(int, t[1:])
            es[key(a, b)] = l
    for line in open(update):
        t = line.split()
        if not t:
            continue
        a, b, l = map(int, t[1:])
        if t[0] == "-e":
            del es[key(a, b)]
        else:
            es[key(a, b)] = l
    with open(out, "w") as f:
        f.write("\n".join(vs) + "\n")
        for (a, b), l in sorted(es.items()):
            f.write(f"e {a} {b} {l}\n")
