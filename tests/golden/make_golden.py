"""Write tests/golden/ref_vectors.npz: input/output vectors of the reference
library (ivan-pi/cbk, src/cbk.cpp and its kernel headers, compiled from its own
sources by oracle/Makefile `ref`) on small seeded cases of every routine.

Run in a container where /root/reference exists and `make -C oracle ref` has
built oracle/_ref/. The fixture lets the oracle be pinned against the
reference on machines that do not have it (the GPU box). Each entry stores the
full argument list (buffers included) and the reference's return status and
output buffers; the FP-contraction build flags are recorded alongside.
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import cbkt  # noqa: E402
import cases  # noqa: E402

OUT = os.path.join(HERE, "ref_vectors.npz")


def main():
    r = cbkt.ref()
    if r is None:
        raise SystemExit("reference not built (make -C oracle ref)")
    arrays = {}
    manifest = {"lib": os.path.basename(r.path), "flags": "-O3 -ffp-contract=fast -fno-math-errno",
                "cases": []}
    for i, (rt, p, opts, dims) in enumerate(cases.sweep(quick=True)):
        # every third trsm / symmetric-solve case keeps the file small; every
        # routine, option set and size class is still represented
        if rt in ("trsm", "posv", "potrs", "sysvnp", "sytrsnp") and i % 3:
            continue
        V, nm = 4, 5
        c = cases.make_case(rt, p, opts, dims, V, nm, seed=5000 + i)
        st, outs = cases.run(r.call, c)
        args = []
        for j, a in enumerate(c["args"]):
            if isinstance(a, np.ndarray):
                key = f"c{i}_a{j}"
                arrays[key] = a
                args.append({"array": key})
            else:
                args.append({"value": a})
        okeys = []
        for j, o in enumerate(outs):
            key = f"c{i}_o{j}"
            arrays[key] = o
            okeys.append(key)
        manifest["cases"].append({"routine": rt, "p": p, "opts": opts, "dims": dims, "V": V,
                                  "nm": nm, "args": args, "outs_idx": c["outs"],
                                  "status": st, "outs": okeys})
    arrays["manifest"] = np.frombuffer(json.dumps(manifest).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **arrays)
    print(OUT, len(manifest["cases"]), "cases", os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
