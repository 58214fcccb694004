"""Writes the golden accepted-index fixtures of tests/golden/ from the CPU oracle.

Each fixture is the serial-RSA output (accepted candidate indices and group labels)
for one model and seed on the Philox candidate stream, as computed by oracle/ (which
restates cuda/braid_cuda.cu's host admission).  The small ones are re-derived by the
oracle in tests as well; the long runs (config 2 to jam, config 3 prefix) exist so the
GPU tests need not rerun a minute-long CPU pack.

    python tests/golden/make_golden.py [name ...]
"""

from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))

import pyoracle  # noqa: E402

# name -> (d, T, terms, seed, pin, candidates, stop_kind)
CASES = {
    **{f"cfg1_d2_T16_s{s}": (2, 16, 16, s, False, 20000, 0) for s in (1, 2, 3, 4)},
    "d2_T32_terms2_s1": (2, 32, 2, 1, False, 50000, 0),
    "d2_T24_terms5_s3": (2, 24, 5, 3, False, 30000, 0),
    "d2_T16_pin_s2": (2, 16, 16, 2, True, 20000, 0),
    "d3_T12_s1": (3, 12, 12, 1, False, 20000, 0),
    "d3_T16_terms4_s5": (3, 16, 4, 5, False, 20000, 0),
    **{f"cfg2_d2_T64_s{s}_jam": (2, 64, 64, s, False, 1 << 40, 1) for s in (1, 2, 3, 4)},
    "cfg3_d3_T128_s1_prefix1M": (3, 128, 128, 1, False, 1 << 20, 0),
    # long scans: 3+1 T=256 (37 head rounds of 7 timesteps in the FP64 test, two slot pairs
    # per lane) and 2+1 T=512 (four slot pairs per lane)
    "d3_T256_s1": (3, 256, 256, 1, False, 6000, 0),
    "d2_T512_s2": (2, 512, 512, 2, False, 8000, 0),
}

# Subpath phase with a windowed stop (these need >= 2e7 subpath candidates, minutes of
# oracle time): name -> (d, T, seed, unique-phase candidates, accept_rate, fill_rate).
# Stored as counts, a digest of the lists and their first/last entries.
# The unique phase is serial RSA over a budget ("budget", the API's UG_STOP_BUDGET) or the
# reference engine's round schedule over it ("ref", what the braid_ug front end runs).
# The full subpath lists and the round records {start, size, survivors, admitted, fresh
# cells} go to a .npz beside the .json.
SUB_CASES = {
    "sub_d2_T16_s1_accept2e-3": (2, 16, 1, 20000, 2e-3, 0.0, "budget"),
    "sub_d2_T16_s2_fill3e-5": (2, 16, 2, 20000, 0.0, 3e-5, "budget"),
    "subref_d2_T16_s1_accept2e-3": (2, 16, 1, 20000, 2e-3, 0.0, "ref"),
}

# Unique phase on the reference engine's round schedule with its --until-accept-rate window
# (rule 2 = braid_cuda, 3 = braid_cuda3d): name -> (d, T, seed, budget, rate).
REF_CASES = {
    "ref_d2_T16_s1_rate2e-6": (2, 16, 1, 1e12, 2e-6),
    "ref_d3_T8_s1_rate1.5e-6": (3, 8, 1, 1e12, 1.5e-6),
    "ref_d2_T64_s1_rate1e-4": (2, 64, 1, 1e12, 1e-4),
    "ref_d3_T32_s2_rate1e-4": (3, 32, 2, 1e12, 1e-4),
}


def list_digest(idx, gid) -> str:
    import hashlib

    import numpy as np

    h = hashlib.sha256()
    h.update(np.asarray(idx, "<i8").tobytes())
    h.update(np.asarray(gid, "<i4").tobytes())
    return h.hexdigest()


def make_sub(name: str) -> dict:
    import numpy as np

    d, T, seed, uniq, ar, fr, rule = SUB_CASES[name]
    m = pyoracle.Model(d, T, seed=seed)
    if rule == "ref":
        r = m.pack_ref(uniq)
        idx, gid, att = r["idx"], r["gid"], r["attempts"]
    else:
        idx, gid, att = m.pack(uniq)
    pos = m.positions_of(idx)
    t0 = time.time()
    sidx, sgid, satt, fresh, rounds = m.pack_sub(pos, gid, first=att, budget=1e15, accept_rate=ar,
                                                 fill_rate=fr, rounds=True)
    dt = time.time() - t0
    all_idx = list(idx) + list(sidx)
    all_gid = list(gid) + list(sgid)
    rec = {
        "name": name, "dim": d, "T": T, "seed": seed, "unique_candidates": uniq, "unique_rule": rule,
        "accept_rate": ar, "fill_rate": fr, "n": int(len(idx)), "nsub": int(len(sidx)),
        "unique_attempts": int(att),
        "attempts": int(att + satt), "sub_attempts": int(satt), "filled_new": int(fresh),
        "digest": list_digest(all_idx, all_gid),
        "head_idx": [int(x) for x in sidx[:64]], "head_gid": [int(x) for x in sgid[:64]],
        "tail_idx": [int(x) for x in sidx[-64:]], "tail_gid": [int(x) for x in sgid[-64:]],
        "oracle_seconds": round(dt, 2),
    }
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(rec, f)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), uidx=np.asarray(idx, np.int64),
                        ugid=np.asarray(gid, np.int32), sidx=np.asarray(sidx, np.int64),
                        sgid=np.asarray(sgid, np.int32), rounds=np.asarray(rounds, np.int64))
    print(f"{name}: N={len(idx)} Nsub={len(sidx)} sub candidates={satt} ({dt:.1f} s)")
    return rec


def make_ref(name: str) -> dict:
    d, T, seed, budget, rate = REF_CASES[name]
    m = pyoracle.Model(d, T, seed=seed)
    t0 = time.time()
    r = m.pack_ref(budget, stop_rate=rate)
    dt = time.time() - t0
    rec = {
        "name": name, "dim": d, "T": T, "seed": seed, "budget": budget, "rate": rate,
        "rule": d, "attempts": int(r["attempts"]), "decided": int(r["decided"]),
        "n": int(len(r["idx"])), "idx": [int(x) for x in r["idx"]],
        "filled": int(r["rounds"][:, 4].sum()),
        "rounds": [[int(v) for v in row] for row in r["rounds"]], "oracle_seconds": round(dt, 2),
    }
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(rec, f)
    print(f"{name}: N={rec['n']} attempts={rec['attempts']} rounds={len(rec['rounds'])} ({dt:.1f} s)")
    return rec


def make(name: str) -> dict:
    d, T, terms, seed, pin, cand, stop = CASES[name]
    m = pyoracle.Model(d, T, terms=terms, seed=seed, pin=pin)
    t0 = time.time()
    idx, gid, att = m.pack(cand, stop_kind=stop)
    dt = time.time() - t0
    rec = {
        "name": name, "dim": d, "T": T, "terms": terms, "seed": seed, "pin_sin1": pin,
        "stop": "fixed_window(batch=2^20, window=2^22, rate=1e-4)" if stop else "budget",
        "candidates": int(att), "n": int(len(idx)), "idx": [int(x) for x in idx],
        "gid": [int(x) for x in gid], "oracle_seconds": round(dt, 2),
    }
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(rec, f)
    print(f"{name}: N={len(idx)} candidates={att} ({dt:.1f} s)")
    return rec


if __name__ == "__main__":
    for n in sys.argv[1:] or list(CASES) + list(SUB_CASES) + list(REF_CASES):
        make_sub(n) if n in SUB_CASES else make_ref(n) if n in REF_CASES else make(n)
