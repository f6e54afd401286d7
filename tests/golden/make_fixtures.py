"""Extract the reference's gas-only golden states into small fixtures.

Reads tests/goldens/<case>/golden.json under /root/reference (the reference's own
regression data: conservative state q1..q5 per region-local cell at the final time,
written by its gas-jump cases, tests/rocpicl_tests/cases.py:443-456) and writes
tests/golden/<name>.json with the case spec and the arrays, unchanged.
Run once in a container that has the reference; the fixtures are committed.
"""
import json
import os
import sys

REF = "/root/reference/tests/goldens"
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {
    "gas_jump_closed_creconst0_np2": "box/gas-jump/closed-creconst0/np2",
    "gas_jump_closed_np2": "box/gas-jump/closed/np2",
    "gas_jump_periodic_np1": "box/gas-jump/periodic/np1",
    "gas_jump_symmetry_np1": "box/gas-jump/symmetry/np1",
}


def main():
    for name, case in CASES.items():
        src = os.path.join(REF, case, "golden.json")
        if not os.path.exists(src):
            print(f"missing {src}", file=sys.stderr)
            continue
        with open(src) as f:
            g = json.load(f)
        out = {
            "source": f"tests/goldens/{case}/golden.json",
            "spec": g["spec"],
            "time": g["state"]["time"],
            "gas": {r: {q: v[q] for q in ("q1", "q2", "q3", "q4", "q5")}
                    for r, v in g["state"]["gas"].items()},
            "totals": {k: v for k, v in g["state"]["totals"].items() if k.startswith("gas_")},
        }
        dst = os.path.join(HERE, name + ".json")
        with open(dst, "w") as f:
            json.dump(out, f, separators=(",", ":"))
        print("wrote", dst)


if __name__ == "__main__":
    main()
