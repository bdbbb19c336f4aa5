"""Writes tests/golden/doc_v7_recon.json: the recon description example of
the reference's usage_docs/doc_v7 (its first ```json block, lines 7-79), as
data for the recon-description tests.  Run in a container that holds the
reference at /root/reference."""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/usage_docs/doc_v7"


def main():
    text = open(SRC).read()
    block = re.search(r"```json\n(.*?)```", text, re.S).group(1)
    cfg = json.loads(block)
    out = os.path.join(HERE, "doc_v7_recon.json")
    with open(out, "w") as f:
        json.dump(cfg, f, indent=4)
        f.write("\n")
    print(out)


if __name__ == "__main__":
    main()
