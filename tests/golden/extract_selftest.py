"""Writes selftest_vectors.json: the reference's 16 self-test cases and their
golden (factorcount, primecount, checksum) triples, read as text from
cl_sieve.cpp:1777-2095 (run_test).  Run once in a container that has the
reference tree mounted; the JSON is the committed fixture."""
import json
import os
import re
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/cl_sieve.cpp"
lines = open(SRC).read().splitlines()
start = next(i for i, l in enumerate(lines) if l.startswith("void run_test("))
cases, cur = [], None
for i in range(start, len(lines)):
    l = lines[i].strip()
    if l.startswith("reset_data("):
        cur = {"modes": 0}
    m = re.match(r"st\.(factorial|primorial|compositorial) = true;", l)
    if m and cur is not None:
        cur["modes"] |= {"factorial": 1, "primorial": 2, "compositorial": 4}[m.group(1)]
    m = re.match(r"st\.(pmin|pmax|nmin|nmax) = (\d+);", l)
    if m and cur is not None:
        cur[m.group(1)] = int(m.group(2))
    m = re.search(r"st\.factorcount == (\d+) && st\.primecount == (\d+) && st\.checksum == 0x([0-9A-Fa-f]+)", l)
    if m and cur is not None:
        cur.update(factorcount=int(m.group(1)), primecount=int(m.group(2)),
                   checksum=m.group(3).upper(), source=f"cl_sieve.cpp:{i + 1}")
        cur["case"] = len(cases) + 1
        cases.append(cur)
        cur = None
assert len(cases) == 16, len(cases)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selftest_vectors.json")
with open(out, "w") as f:
    json.dump(cases, f, indent=1)
print(out, len(cases))
