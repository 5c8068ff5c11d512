"""Writes oracle_digests.json: SHA-256 digests of the oracle's colour and depth
buffers and its counters for every scene of scene_cases.golden_cases().

Run from the repository root after building the oracle (make -C oracle):
    python tests/golden/make_golden.py              # every scene
    python tests/golden/make_golden.py NAME ...     # only these (others kept)
    python tests/golden/make_golden.py --missing    # only scenes without a digest
    python tests/golden/make_golden.py --jobs 8 ... # oracle runs in 8 processes
The digests pin the oracle and the seeded scene generators against drift; the
GPU parity tests compare the HIP path with them as well as with the live oracle.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import refcpu  # noqa: E402
from parity import digest  # noqa: E402
from scene_cases import golden_cases  # noqa: E402


_BUILDERS = {}


def _one(name):
    scene = _BUILDERS[name]()
    rgba, depth, st = refcpu.draw(scene)
    return name, digest(rgba, depth, st)


def main():
    path = os.path.join(HERE, "oracle_digests.json")
    args = sys.argv[1:]
    jobs = 1
    if args[:1] == ["--jobs"]:
        jobs = int(args[1])
        args = args[2:]
    old = json.load(open(path)) if args and os.path.exists(path) else {}
    cases = golden_cases()
    names = {n for n, _ in cases}
    if args == ["--missing"]:
        todo = names - set(old)
    elif args:
        todo = set(args)
        unknown = todo - names
        if unknown:
            raise SystemExit(f"unknown scenes: {sorted(unknown)}")
    else:
        todo = names
    out = {n: d for n, d in old.items() if n in names}
    _BUILDERS.update(cases)
    order = [n for n, _ in cases if n in todo]
    if jobs > 1:
        import multiprocessing
        with multiprocessing.get_context("fork").Pool(jobs) as pool:
            results = pool.imap_unordered(_one, order)
            for name, d in results:
                out[name] = d
                print(name, d["triangles"], d["pixels_shaded"], flush=True)
    else:
        for name in order:
            name, out[name] = _one(name)
            print(name, out[name]["triangles"], out[name]["pixels_shaded"], flush=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
