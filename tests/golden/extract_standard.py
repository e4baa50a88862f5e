"""Extract the golden rows of the reference's standard table used to pin the oracle.

Reads test/verification/standard_reference.csv of the reference (a data file of its own tests,
written by scripts/standard/update_reference.jl) and keeps the Newton/LU columns whose solve the
oracle restates: "Newton Static LU", "Newton Backtracking LU", "Newton reuse 3 Backtracking LU"
and "Newton DogLeg LU". Run from the repository root:
python tests/golden/extract_standard.py /root/reference
"""
import csv
import sys
from pathlib import Path

COLUMNS = ("Newton Static LU", "Newton Backtracking LU", "Newton reuse 3 Backtracking LU", "Newton DogLeg LU")


def main(ref_root: str) -> None:
    src = Path(ref_root) / "test" / "verification" / "standard_reference.csv"
    dst = Path(__file__).with_name("standard_reference_newton_lu.csv")
    with open(src, newline="") as f:
        rows = [r for r in csv.DictReader(f) if r["column"] in COLUMNS]
    with open(dst, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    print(f"{len(rows)} rows -> {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
