"""Collect the reference's own test data into small fixtures under tests/golden.

Run in a checkout that has the reference mounted (default /root/reference).
Only data is copied: the example puzzle definitions, the reference's expected
output files (results.bin, stats tables, solution lists, capability-profile
plots) and the expected stdout strings of each reference test (test.info
"expected_stdout_string" lines, stored as JSON).  No reference source is
copied.
"""
import json
import os
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    os.makedirs(os.path.join(OUT, "puzzles"), exist_ok=True)
    for n in ("3_3", "4_4", "5_5", "5_5_bis"):
        src = os.path.join(REF, "doc", "example", n + ".txt")
        if os.path.exists(src):
            shutil.copyfile(src, os.path.join(OUT, "puzzles", n + ".txt"))
    copies = []
    for d in sorted(os.listdir(os.path.join(REF, "tests"))) + ["../tests_CUDA/" + d for d in sorted(os.listdir(os.path.join(REF, "tests_CUDA")))]:
        base = os.path.join(REF, "tests", d)
        refdir = os.path.join(base, "references")
        if os.path.isdir(refdir):
            for f in sorted(os.listdir(refdir)):
                if f.endswith("_vtk.txt") or "_level_" in f or f.endswith("_glutton.txt"):
                    continue
                name = os.path.basename(d) + "__" + f
                shutil.copyfile(os.path.join(refdir, f), os.path.join(OUT, name))
                copies.append(name)
    expected = {}
    for d in sorted(os.listdir(os.path.join(REF, "tests"))) + ["../tests_CUDA/" + d for d in sorted(os.listdir(os.path.join(REF, "tests_CUDA")))]:
        info = os.path.join(REF, "tests", d, "test.info")
        if not os.path.exists(info):
            continue
        lines = open(info).read().splitlines()
        entry = {"strings": []}
        for line in lines:
            if line.startswith("args:"):
                parts = line[len("args:"):].split()
                entry["puzzle"] = os.path.basename(parts[0]).replace(".txt", "")
                entry["feature"] = parts[2]
            elif line.startswith("expected_stdout_string:"):
                entry["strings"].append(line[len("expected_stdout_string:"):])
        expected[os.path.basename(d)] = entry
    with open(os.path.join(OUT, "expected_stdout.json"), "w") as f:
        json.dump(expected, f, indent=1, sort_keys=True)
    print("copied", len(copies), "files; expected strings for", len(expected), "tests")
    collect_situation_profiles()
    collect_profiles()


def collect_profiles():
    """tests/profile_<n>: the profile feature's six VTK histograms (per-level
    extrema of the capability profiles) under profile/<n>/."""
    root = os.path.join(OUT, "profile")
    for d in sorted(os.listdir(os.path.join(REF, "tests"))):
        if not d.startswith("profile_"):
            continue
        n = d[len("profile_"):]
        os.makedirs(os.path.join(root, n), exist_ok=True)
        refdir = os.path.join(REF, "tests", d, "references")
        for f in sorted(os.listdir(refdir)):
            if f.endswith("_vtk.txt"):
                shutil.copyfile(os.path.join(refdir, f), os.path.join(root, n, f))
    print("profiles:", ", ".join(sorted(os.listdir(root))))


def collect_situation_profiles():
    """tests/situation_profile_<n>: every reference output file (VTK line
    plots and surfaces of the capability profiles) under
    situation_profile/<n>/, and the --initial_situation argument of each test
    in situation_profile/situations.json."""
    situations = {}
    root = os.path.join(OUT, "situation_profile")
    for d in sorted(os.listdir(os.path.join(REF, "tests"))):
        if not d.startswith("situation_profile_"):
            continue
        n = d[len("situation_profile_"):]
        os.makedirs(os.path.join(root, n), exist_ok=True)
        refdir = os.path.join(REF, "tests", d, "references")
        for f in sorted(os.listdir(refdir)):
            shutil.copyfile(os.path.join(refdir, f), os.path.join(root, n, f))
        for line in open(os.path.join(REF, "tests", d, "test.info")).read().splitlines():
            if line.startswith("args:") and "--initial_situation=" in line:
                situations[n] = line.split('--initial_situation="', 1)[1].split('"', 1)[0]
    with open(os.path.join(root, "situations.json"), "w") as f:
        json.dump(situations, f, indent=1, sort_keys=True)
    print("situation profiles:", ", ".join(sorted(situations)))


if __name__ == "__main__":
    main()
