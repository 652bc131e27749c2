"""Writes blake3_vectors.json: published BLAKE3 test vectors.

The values are the official BLAKE3 test vectors (BLAKE3 repository,
test_vectors/test_vectors.json, "hash" field, default 32-byte output) for
inputs of `input_len` bytes where byte i = i % 251, plus the two plain
examples BLAKE3("") and BLAKE3("abc").  They are transcribed constants: this
script only records them with their provenance; it computes nothing.
"""
import json
import os

VECTORS = {
    "0": "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262",
    "1": "2d3adedff11b61f14c886e35afa036736dcd87a74d27b5c1510225d0f592e213",
    "1023": "10108970eeda3eb932baac1428c7a2163b0e924c9a9e25b35bba72b28f70bd11",
    "1024": "42214739f095a406f3fc83deb889744ac00df831c10daa55189b5d121c855af7",
    "1025": "d00278ae47eb27b34faecf67b4fe263f82d5412916c1ffd97c8cb7fb814b8444",
    "2048": "e776b6028c7cd22a4d0ba182a8bf62205d2ef576467e838ed6f2529b85fba24a",
    "2049": "5f4d72f40d7a5f82b15ca2b2e44b1de3c2ef86c426c95c1af0b6879522563030",
    "3072": "b98cb0ff3623be03326b373de6b9095218513e64f1ee2edd2525c7ad1e5cffd2",
}
PLAIN = {"abc": "6437b3ac38465133ffb63b75273a8db548c558465d79db03fd359c6cd5bd9d85"}

if __name__ == "__main__":
    out = {
        "source": "BLAKE3 official test vectors (input byte i = i % 251), hash mode, 32-byte output",
        "pattern_inputs": VECTORS,
        "plain_inputs": PLAIN,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blake3_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(path)
