"""Extract known-answer vectors from the reference's own test sources.

Run in a container where the reference is mounted (default /root/reference):

    python tests/golden/extract_reference_vectors.py [/path/to/reference]

It reads only the DATA that the reference's GoogleTest files hold (array and
scalar literals inside each TEST block) and writes them as JSON fixtures next
to this script:

  ntt_vectors.json          test/operation/ntt.cpp          (NTT / INTT KATs)
  montgomery_vectors.json   test/montgomery/unit.cpp         (mont_mult KATs)
  precomputed_tiers.json    test/ckks/precomputed/default.cpp (prime chains,
                            per-prime constants, twiddle prefixes, key-switch
                            and rescale constants for every tier)

No reference code is executed or copied; the fixtures are literal values.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

NUM = r"-?\d+"
BLOCK_RE = re.compile(r"^TEST\((\w+),\s*(\w+)\)\s*\{", re.M)
ARRAY_RE = re.compile(
    r"(?:u?int64_t|int)\s+(\w+)\s*\[\s*\w*\s*\]\s*(?:=\s*)?\{([^{}]*)\}", re.S)
SCALAR_RE = re.compile(r"(?:u?int64_t|int|auto)\s+(\w+)\s*=\s*(" + NUM + r")\s*;")
VECTOR_RE = re.compile(
    r"std::vector<u?int64_t>\s+(\w+)\s*=\s*\{([^{}]*)\}\s*;", re.S)
NESTED_RE = re.compile(
    r"std::vector<std::vector<u?int64_t>>\s+(\w+)\s*=\s*\{(.*?)\}\s*;", re.S)
EXPECT_RE = re.compile(r"EXPECT_EQ\(precomputed\.(\w+),\s*(" + NUM + r")\)")
CALL_RE = re.compile(r"desilo::fhe::(\w+)\(")


def ints(text):
    return [int(v) for v in re.findall(NUM, text)]


def blocks(source):
    heads = list(BLOCK_RE.finditer(source))
    for i, m in enumerate(heads):
        end = heads[i + 1].start() if i + 1 < len(heads) else len(source)
        yield m.group(1), m.group(2), source[m.end():end]


def parse_block(body):
    out = {"arrays": {}, "scalars": {}, "calls": CALL_RE.findall(body)}
    for m in ARRAY_RE.finditer(body):
        out["arrays"][m.group(1)] = ints(m.group(2))
    for m in VECTOR_RE.finditer(body):
        out["arrays"][m.group(1)] = ints(m.group(2))
    for m in NESTED_RE.finditer(body):
        out["arrays"][m.group(1)] = [ints(g) for g in
                                     re.findall(r"\{([^{}]*)\}", m.group(2))]
    for m in SCALAR_RE.finditer(body):
        out["scalars"][m.group(1)] = int(m.group(2))
    for m in EXPECT_RE.finditer(body):
        out["scalars"][m.group(1)] = int(m.group(2))
    tier = re.search(r'from_tier\("(\w+)"\)', body)
    if tier:
        out["tier"] = tier.group(1)
    if "for_bootstrap()" in body:
        out["tier"] = "bootstrap_gold"
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    jobs = {
        "ntt_vectors.json": "test/operation/ntt.cpp",
        "montgomery_vectors.json": "test/montgomery/unit.cpp",
        "precomputed_tiers.json": "test/ckks/precomputed/default.cpp",
    }
    for name, rel in jobs.items():
        with open(os.path.join(ref, rel)) as f:
            source = f.read()
        fixture = {"source": rel, "tests": {}}
        for suite, test, body in blocks(source):
            fixture["tests"][test] = parse_block(body)
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fixture, f, indent=1)
        print(f"{name}: {len(fixture['tests'])} tests from {rel}")


if __name__ == "__main__":
    main()
