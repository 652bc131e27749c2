"""Writes phi8_table.json: the 256 values of the reference's phi8 embedding
table (PHI_8_TABLE, crates/flock-core/src/field/phi8.rs:19-1044), read as
numeric data from the reference source.  Only runs where /root/reference
exists; the JSON it writes is the committed fixture.
"""
import json
import os
import re

SRC = "/root/reference/crates/flock-core/src/field/phi8.rs"

if __name__ == "__main__":
    text = open(SRC).read()
    body = text[text.index("PHI_8_TABLE"):text.index("];")]
    pairs = re.findall(r"lo:\s*0x([0-9a-fA-F]+),\s*hi:\s*0x([0-9a-fA-F]+)", body)
    assert len(pairs) == 256, len(pairs)
    table = [[lo.lower(), hi.lower()] for lo, hi in pairs]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "phi8_table.json")
    with open(path, "w") as f:
        json.dump({"source": "crates/flock-core/src/field/phi8.rs:19-1044 PHI_8_TABLE (lo, hi hex)",
                   "table": table}, f, indent=0)
    print(path, len(table))
