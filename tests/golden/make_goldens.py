"""Write the golden fixtures in this folder from the reference build.

Runs oracle/_ref/ref_driver (the unmodified reference sources compiled by
oracle/Makefile) where the reference sources are present; the committed
fixtures are what every other machine reads.

    python tests/golden/make_goldens.py [--quick]

Outputs (all JSON, committed):
  reference_counts.json     counts + schedule dumps + CSR hashes on wiki-vote
                            and the R-MAT scale-13 graph
  reference_schedules.json  Schedule dumps for every connected pattern of 3-6
                            vertices (and the named 7-vertex ones) under
                            several planner settings and graph statistics
  reference_motifs.json     MotifGenerator output for 2..6 vertices
  wiki-vote_input.gz        the reference's dataset/wiki-vote_input (data)
"""
import gzip
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import graphpi_amd as gp  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
WIKI = "/root/reference/dataset/wiki-vote_input"
THREADS = "8"


def ref(args, timeout=600):
    out = subprocess.run([REF] + args, capture_output=True, text=True, timeout=timeout).stdout
    line = [l for l in out.splitlines() if l.startswith("JSON ")]
    if not line:
        raise RuntimeError("reference produced no JSON for " + " ".join(args))
    return json.loads(line[0][5:])


def adj_str(a):
    return "".join("1" if x else "0" for x in a.flatten())


COUNT_PATTERNS = {
    "wiki-vote": ["triangle", "rectangle", "qg3", "house", "4clique", "5clique", "pentagon", "wedge",
                  "tri_edge_tri"],
    "rmat13": ["triangle", "rectangle", "house", "qg3", "4clique", "5clique", "pentagon", "wedge",
               "tri_edge_tri", "hourglass", "cycle_6_tri"],
}
EXTRA_SETTINGS = {  # (pm, rt, iep) beyond (1, 1, 1), on the small patterns
    "rmat13": [("triangle", 0, 0, 0), ("rectangle", 0, 0, 0), ("house", 0, 1, 0), ("rectangle", 0, 1, 0),
               ("house", 1, 1, 0), ("pentagon", 1, 1, 0), ("4clique", 0, 2, 0), ("house", 2, 2, 0),
               ("house", 3, 1, 1), ("qg3", 0, 0, 0)],
}


def main():
    quick = "--quick" in sys.argv
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    with open(WIKI, "rb") as f, gzip.open(os.path.join(HERE, "wiki-vote_input.gz"), "wb", 9) as g:
        g.write(f.read())

    tmp = tempfile.mkdtemp()
    rmat_path = os.path.join(tmp, "rmat13.txt")
    h = gp.HostGraph.rmat(13, 12, 0.57, 0.19, 0.19, 1)
    h.write_text(rmat_path)
    graphs = {"wiki-vote": WIKI, "rmat13": rmat_path}

    counts = []
    for gname, path in graphs.items():
        jobs = [(p, 1, 1, 1) for p in COUNT_PATTERNS[gname]] + EXTRA_SETTINGS.get(gname, [])
        if quick:
            jobs = jobs[:3]
        for pname, pm, rt, iep in jobs:
            a = gp.PATTERNS[pname]
            n = a.shape[0]
            try:
                r = ref(["count", path, str(n), adj_str(a), str(pm), str(rt), str(iep), THREADS], timeout=900)
            except subprocess.TimeoutExpired:
                print("timeout", gname, pname, pm, rt, iep, flush=True)
                continue
            r.update({"graph": gname, "pattern": pname, "pattern_adj": adj_str(a), "pm": pm, "rt": rt,
                      "iep": iep})
            print(gname, pname, pm, rt, iep, r["count"], "%.2fs" % r["time_count"], flush=True)
            counts.append(r)
    with open(os.path.join(HERE, "reference_counts.json"), "w") as f:
        json.dump(counts, f, indent=0)

    # graph statistics the planner sees (v_cnt, e_cnt, tri_cnt)
    stats = []
    for r in counts:
        if r["pattern"] == "triangle" and r["pm"] == 1:
            stats.append((r["graph"], r["v_cnt"], r["e_cnt"], r["tri_cnt"]))
    stats.append(("rmat22", 2009103, 65241502, 0))  # tri_cnt unknown here; the model still runs
    settings = [(1, 1, 1), (1, 1, 0), (0, 1, 0), (0, 0, 0), (0, 1, 1), (2, 2, 0), (3, 1, 1), (1, 2, 1),
                (0, 2, 0), (2, 1, 1)]
    pats = []
    for n in range(3, 7 if not quick else 5):
        pats += [(n, adj_str(a)) for a in gp.motifs(n)]
    for name in ("clique_7_minus",):
        a = gp.PATTERNS[name]
        pats.append((7, adj_str(a)))
    scheds = []
    for gname, v, e, t in stats:
        for n, s in pats:
            for pm, rt, iep in settings:
                if n >= 6 and (pm, rt, iep) not in ((1, 1, 1), (0, 1, 0), (2, 2, 0)):
                    continue
                r = ref(["sched", str(n), s, str(pm), str(rt), str(iep), str(v), str(e), str(t)])
                scheds.append({"graph": gname, "size": n, "adj": s, "pm": pm, "rt": rt, "iep": iep,
                               "v_cnt": v, "e_cnt": e, "tri_cnt": t, "valid": r["valid"],
                               "schedule": r["schedule"]})
    with open(os.path.join(HERE, "reference_schedules.json"), "w") as f:
        json.dump(scheds, f, separators=(",", ":"))
    print("schedules", len(scheds))

    motifs = {}
    for n in range(2, 7):
        motifs[str(n)] = ref(["motif", str(n)])["patterns"]
    with open(os.path.join(HERE, "reference_motifs.json"), "w") as f:
        json.dump(motifs, f)
    shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
