"""Regenerate csrc/ligerito_ladders.inc from the reference's Ligerito security
ladders (crates/flock-core/configs/ligerito/m{22..35}_{fast,slim,secure}.toml,
embedded by CF/pcs/ligerito.rs:256-345 and mapped to a ProverConfig by
LigeritoSecurityConfig::to_prover_verifier_configs, :963-1013).

The ladders are protocol parameters (rates, column splits, query counts and
grinding bits fixed by a soundness analysis), so they are carried as data: one
C++ initializer row per (m, profile).  Only the fields the prover and verifier
read are kept; `hash` is overridden by PcsParams.merkle_hash
(CF/pcs/commit.rs:96-104), so it is not carried.

Usage: python gen_ligerito_ladders.py [CONFIG_DIR]
(default: /root/reference/crates/flock-core/configs/ligerito).
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "csrc", "ligerito_ladders.inc")
PROFILES = ("fast", "slim", "secure")
LEVEL_KEYS = ("log_inv_rate", "log_msg_cols", "log_num_interleaved", "k_recursive", "queries",
              "grinding_bits", "fold_grinding_bits", "ood_samples")


def parse(text):
    """The subset of TOML these files use: `key = value` lines, [[levels]] and
    [final_block] tables; integer values are kept, the rest ignored."""
    top, levels, final, cur = {}, [], {}, None
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if line == "[[levels]]":
            cur = {}
            levels.append(cur)
            continue
        if line == "[final_block]":
            cur = final
            continue
        m = re.fullmatch(r"([A-Za-z_0-9]+)\s*=\s*(.+)", line)
        if not m:
            raise ValueError(f"unparsed line: {raw!r}")
        key, val = m.group(1), m.group(2).strip()
        if re.fullmatch(r"-?\d+", val):
            val = int(val)
        (top if cur is None else cur)[key] = val
    return top, levels, final


def main():
    cfg_dir = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/crates/flock-core/configs/ligerito"
    rows = []
    for m in range(22, 36):
        for p, prof in enumerate(PROFILES):
            top, levels, final = parse(open(os.path.join(cfg_dir, f"m{m}_{prof}.toml")).read())
            assert top["m"] == m and top["log_n"] == m - 7, (m, prof)
            assert len(levels) <= 8
            lv = ", ".join("{" + ", ".join(str(l[k]) for k in LEVEL_KEYS) + "}" for l in levels)
            rows.append(f"    {{{m}, {p}, {top['log_n']}, {top['initial_k']}, {final['yr_log_n']}, "
                        f"{len(levels)}, {{{lv}}}}},")
    with open(OUT, "w") as f:
        f.write("// Generated by gen_ligerito_ladders.py from the reference's Ligerito security\n"
                "// ladders (crates/flock-core/configs/ligerito/*.toml).  Row: {m, profile (0 fast,\n"
                "// 1 slim, 2 secure), log_n, initial_k, yr_log_n, n_levels, levels {log_inv_rate,\n"
                "// log_msg_cols, log_num_interleaved, k_recursive, queries, grinding_bits,\n"
                "// fold_grinding_bits, ood_samples}}.\n")
        f.write("\n".join(rows) + "\n")
    print(OUT, len(rows), "ladders")


if __name__ == "__main__":
    main()
