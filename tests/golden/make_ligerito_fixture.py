"""Writes ligerito_ladders.json: the reference's embedded Johnson ladders
(vendor/flock-mod/crates/flock-core/configs/ligerito/m{22..35}_{fast,slim,slim3}.toml),
reduced to the fields a ladder solver must reproduce. Run in a container
that holds the reference tree; the JSON is the committed fixture.
"""
import json
import os
import sys

import tomli

SRC = "/root/reference/vendor/flock-mod/crates/flock-core/configs/ligerito"
PROFILES = {"fast": 1, "slim": 2, "slim3": 3}  # L0 inverse-rate exponent
LEVEL_KEYS = ("log_inv_rate", "log_msg_cols", "log_num_interleaved", "k_recursive", "regime", "eta",
              "queries", "grinding_bits", "fold_grinding_bits", "ood_samples", "target_security_bits",
              "expected_eps_pg_bits", "expected_eps_query_bits", "expected_eps_ood_bits")


def main(out: str) -> None:
    rows = []
    for m in range(22, 36):
        for name, rate in PROFILES.items():
            with open(os.path.join(SRC, f"m{m}_{name}.toml"), "rb") as f:
                cfg = tomli.load(f)
            rows.append(dict(
                profile=name, log_inv_rate=rate, m=cfg["m"], log_n=cfg["log_n"], initial_k=cfg["initial_k"],
                target_security_bits=cfg["target_security_bits"], analysis_version=cfg["analysis_version"],
                field=cfg["field"], grinding_step=cfg["grinding_step"], yr_log_n=cfg["final_block"]["yr_log_n"],
                levels=[{k: lv.get(k) for k in LEVEL_KEYS} for lv in cfg["levels"]]))
    with open(out, "w") as f:
        f.write("[\n" + ",\n".join(json.dumps(r, sort_keys=True, separators=(",", ":")) for r in rows) + "\n]\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "ligerito_ladders.json"))
