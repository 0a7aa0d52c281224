"""A FrameworkBatch laid out the way the reference stores it (src/batch.jl:72-104), with every
table taken from the FP64 oracle: the input pa_batch_from_tables takes over from a Julia host.

Distinct frameworks are stored once (by object identity here), atoms sorted by grid cell with a
stable counting sort (src/cell.jl:96-118; the order only changes summation order), LJ types
compacted to those present in the frameworks and the guest (src/batch.jl:305-317), the coupled
or full k table with the host structure factor, constant_offset, the self-term half-range, the
hard-core bound B_s and K_min(site, type) (src/batch.jl:395,409-423)."""
from __future__ import annotations

import numpy as np

import pureadsorb_amd as pa
from conftest import oracle_for


def reference_tables(fws, ff, guest, ewald, *, fullk=False, cell_edge=4.0):
    uniq, fw_of = [], []
    for fw in fws:
        for u, v in enumerate(uniq):
            if v is fw:
                fw_of.append(u)
                break
        else:
            uniq.append(fw)
            fw_of.append(len(uniq) - 1)
    orig_types = [[pa.typeindex(ff, s + "_") for s in fw.symbols] for fw in uniq]
    present = sorted(set(t for ts in orig_types for t in ts) | set(int(t) for t in guest.types))
    c_of = {t: c for c, t in enumerate(present)}
    ntc = len(present)
    pos, types, charges, cells, alphas, ks, kpref, S = [], [], [], [], [], [], [], []
    atom_off, k_off = [0], [0]
    const, half, bs, kmin = [], [], [], []
    for fw, ts in zip(uniq, orig_types):
        o = oracle_for(fw, ff, guest, ewald)
        cell = np.asarray(fw.cell, dtype=np.float64)
        frac = np.asarray(fw.frac, dtype=np.float64) % 1.0
        # stable counting sort by grid cell, i fastest
        nc = np.maximum(1, np.floor(np.asarray(pa.perpendicular_lengths(cell)) / cell_edge)).astype(int)
        ci = np.minimum((frac * nc).astype(int), nc - 1)
        key = ci[:, 0] + nc[0] * (ci[:, 1] + nc[1] * ci[:, 2])
        order = np.argsort(key, kind="stable")
        cart = np.asarray(fw.frac, dtype=np.float64) @ cell.T
        pos.append(cart[order])
        types.append(np.array([c_of[t] for t in ts], dtype=np.int32)[order])
        charges.append(np.asarray(fw.charges, dtype=np.float64)[order])
        atom_off.append(atom_off[-1] + fw.natoms)
        cells.append(cell)
        alphas.append(o.alpha)
        kv, kp, sk, _ = o.ktable(full=fullk)
        charged = bool(np.any(np.asarray(guest.charges) != 0.0))
        if not charged:
            kv, kp, sk = kv[:0], kp[:0], sk[:0]
        ks.append(kv)
        kpref.append(kp)
        S.append(sk)
        k_off.append(k_off[-1] + len(kp))
        const.append(o.constant_offset)
        half.append(o.self_halfrange)
        bs.append(o.hardcore_bound())
        q = np.asarray(fw.charges, dtype=np.float64)
        km = np.zeros((len(guest.charges), ntc))
        for a, qa in enumerate(guest.charges):
            K = pa.KE * qa * q
            for j, t in enumerate(ts):
                km[a, c_of[t]] = min(km[a, c_of[t]], K[j])
        kmin.append(km)
    idx = np.array(present)
    return {
        "framework_of": np.array(fw_of, dtype=np.int32), "atom_offsets": np.array(atom_off, dtype=np.int32),
        "positions": np.concatenate(pos), "types": np.concatenate(types), "charges": np.concatenate(charges),
        "cells": np.array(cells), "alphas": np.array(alphas), "k_offsets": np.array(k_off, dtype=np.int32),
        "ks": np.concatenate(ks).reshape(-1, 3), "kprefactor": np.concatenate(kpref), "Shost": np.concatenate(S),
        "constant_offset": np.array(const), "self_term_halfrange": np.array(half),
        "sigma": np.asarray(ff.sigma)[np.ix_(idx, idx)], "epsilon": np.asarray(ff.epsilon)[np.ix_(idx, idx)],
        "compact_to_orig": idx.astype(np.int32), "bs": np.array(bs), "kmin": np.array(kmin),
        "cutoff": ff.cutoff, "ewald_cutoff": ewald.cutoff, "fullk": fullk,
    }
