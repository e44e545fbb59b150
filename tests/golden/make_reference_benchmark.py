"""Write reference_benchmark.json: the solver-scaling records the reference publishes.

Source (data, not code): gradient-domain/docs/benchmark.json of
Giankaqui/visual-computing, produced by ``gradient-domain benchmark`` with the
default seed 0 and tolerance 1e-8 (cli.py:164, 198-204; benchmark.py:98-149).
Only the fields the oracle is pinned on are kept.
"""

import json
from pathlib import Path

SRC = Path("/root/reference/gradient-domain/docs/benchmark.json")
OUT = Path(__file__).resolve().parent / "reference_benchmark.json"


def main():
    data = json.loads(SRC.read_text())
    keep = {}
    for method, records in data.items():
        keep[method] = [
            {k: r[k] for k in ("shape", "iterations", "relative_residual", "max_error", "seconds")}
            for r in records
        ]
    OUT.write_text(json.dumps({"source": "gradient-domain/docs/benchmark.json", "seed": 0,
                               "tolerance": 1e-8, "records": keep}, indent=1) + "\n")


if __name__ == "__main__":
    main()
