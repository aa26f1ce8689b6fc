"""Writes example_ltlbs.txt: ExampleLTLF.json in LTL_BS's text trace format
(readFile, LTL_BS.cu:1535-1607), with the APs a0, a1 renamed to the letters
a, b (the text format names APs by single letters).  The alphabet line lists
the letters in reverse order to exercise the format's per-line value order.

    python tests/golden/make_text_fixture.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    j = json.load(open(os.path.join(HERE, "ExampleLTLF.json")))
    names = j["atomic_propositions"]
    letters = [chr(ord("a") + k) for k in range(len(names))]
    order = list(reversed(range(len(names))))      # values listed as b, a

    def line(tr):
        n = len(tr[names[0]])
        return ";".join(",".join(str(int(tr[names[a]][k])) for a in order) for k in range(n))

    out = [line(tr) for tr in j["positive_traces"]] + ["---"] + [line(tr) for tr in j["negative_traces"]]
    out += ["---", "".join(letters[a] for a in order)]
    with open(os.path.join(HERE, "example_ltlbs.txt"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
