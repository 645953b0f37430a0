"""Writes tests/golden/ref_netgen.json from the reference's own NetGen / ConstGen.

Runs oracle/_ref/ref_netgen (built by oracle/build_ref.sh from the reference's
macro.cpp, model/model.cpp and model/netgen.cpp) on the specs below and stores,
per spec, the synapse count, the neuron types/in/out degrees digests, SHA-256
digests of the synapse arrays (src, tar, weight, delay as raw little-endian
bytes), the first synapses in clear, and the ConstGen arrays.  The fixture is
data (inputs and outputs of the reference), small enough to commit.
"""
import hashlib
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402

SPECS = {
    "brunel_20k_clock777": ["timestep 0.0001", "clock 777", "srand 1", "brunel 20000"],
    "brunel_200k_clock123456789012": ["timestep 0.0001", "clock 123456789012", "srand 1", "brunel 200000"],
    "vogel_50k_clock5": ["timestep 0.0001", "clock 5", "srand 3", "vogel 50000"],
    "ranges_all_connectors": ["timestep 0.0001", "dw 0.01", "clock 99", "srand 7", "pop 300 2 1", "pop 200 5 0",
                              "pop 50 5 0", "proj 0 1 0.001 0.05 0.001 0.002 0.2",
                              "proj 1 2 -0.05 -0.01 0.0005 0.0015 1", "proj 2 2 0.1 0.1 0.001 0.003 0",
                              "proj 1 1 0.02 0.02 0.0012 0.0012 0.05"],
    "shuffle_quota": ["timestep 0.0001", "clock 4", "srand 11", "shuffle 1", "pop 120 2 1", "pop 90 5 0",
                      "proj 0 1 0.001 0.004 0.001 0.002 0.3", "proj 1 1 -0.002 -0.002 0.0014 0.0014 0.1"],
    "duplicate_projection_quirk": ["timestep 0.0001", "clock 8", "srand 2", "pop 60 2 1", "pop 40 5 0",
                                   "proj 0 1 0.001 0.001 0.001 0.001 0.5", "proj 0 1 0.002 0.002 0.001 0.001 0.5",
                                   "proj 1 1 0.003 0.003 0.001 0.001 0.2", "proj 1 1 0.004 0.004 0.002 0.002 0.2"],
    "timestep_1ms": ["timestep 0.001", "clock 31", "srand 1", "brunel 5000"],
    # small projections where the 1.1p acceptance runs out before gen_num: the reference still
    # emits gen_num synapses (the unfilled genlist entries are first source -> first target)
    "quota_tail_single": ["timestep 0.0001", "dw 0.01", "clock 12", "srand 1", "pop 10 2 1", "pop 10 5 0",
                          "proj 0 1 0.001 0.001 0.001 0.001 0.2"],
    "quota_tail_chain_clock27": ["timestep 0.0001", "dw 0.002", "clock 27", "srand 5", "pop 10 2 1", "pop 12 5 0",
                                 "pop 7 5 0", "proj 0 1 0.001 0.009 0.001 0.002 0.2",
                                 "proj 1 2 -0.02 -0.01 0.0005 0.0015 0.15", "proj 2 1 0.003 0.011 0.001 0.003 0.3"],
    "quota_tail_chain_clock33": ["timestep 0.0001", "dw 0.002", "clock 33", "srand 5", "pop 10 2 1", "pop 12 5 0",
                                 "pop 7 5 0", "proj 0 1 0.001 0.009 0.001 0.002 0.2",
                                 "proj 1 2 -0.02 -0.01 0.0005 0.0015 0.15", "proj 2 1 0.003 0.011 0.001 0.003 0.3"],
}


def digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def main():
    if not po.ref_netgen_available():
        sys.exit("oracle/_ref/ref_netgen missing: run oracle/build_ref.sh (needs the reference sources)")
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, spec in SPECS.items():
            r = po.run_ref_netgen(spec, d)
            out[name] = {
                "spec": spec,
                "n": r["n"],
                "n_syn": int(r["src"].size),
                "types": digest(r["types"]),
                "nxt": digest(r["nxt"]),
                "pre": digest(r["pre"]),
                "src": digest(r["src"]),
                "tar": digest(r["tar"]),
                "weight": digest(r["weight"]),
                "delay": digest(r["delay"]),
                "head": [[int(a), int(b), float(c), float(e)] for a, b, c, e in
                         zip(r["src"][:8], r["tar"][:8], r["weight"][:8], r["delay"][:8])],
                "nlifconst": r["nlifconst"],
                # the reference leaves Network::lifconst uninitialised unless a model sets it
                "lifconst": [float(x) for x in r["lifconst"]] if r["nlifconst"] else None,
                "life1": [float(x) for x in r["life1"]],
                "life2": [float(x) for x in r["life2"]],
                "life3": [float(x) for x in r["life3"]],
                "stdp1": [float(x) for x in r["stdp1"]],
            }
    with open(os.path.join(HERE, "ref_netgen.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(out), "fixtures")


if __name__ == "__main__":
    main()
