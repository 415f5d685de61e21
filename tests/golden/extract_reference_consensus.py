"""Writes reference_consensus.json: the expected consensus strings of the
reference's 18 end-to-end tests (test/spoa_test.cpp:154-440), keyed by
test name with the scoring mode each test sets up.  Run in the build
container only (it reads /root/reference); the JSON is the fixture."""
import json
import os
import re

SRC = "/root/reference/test/spoa_test.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_consensus.json")
TYPES = {"kSW": 0, "kNW": 1, "kOV": 2}


def main():
    text = open(SRC).read()
    out = {}
    for m in re.finditer(r"TEST_F\(SpoaTest, (\w+)\) \{(.*?)\n\}", text, re.S):
        name, body = m.group(1), m.group(2)
        setup = re.search(r"Setup\(AlignmentType::(\w+), ([-\d]+), ([-\d]+), ([-\d]+), ([-\d]+), "
                          r"([-\d]+), ([-\d]+), (true|false)\)", body)
        cons = re.search(r"std::string c =(.*?);", body, re.S)
        if not setup or not cons:
            continue
        params = [TYPES[setup.group(1)]] + [int(setup.group(k)) for k in range(2, 8)]
        consensus = "".join(re.findall(r'"([ACGT]*)"', cons.group(1)))
        out[name] = {"params": params, "quality": setup.group(8) == "true", "consensus": consensus}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(len(out), "tests ->", OUT)


if __name__ == "__main__":
    main()
