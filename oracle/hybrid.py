"""CPU restatement of the reference's cost order and hybrid CPU/GPU split.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
gc_amd_cost_slots, gc_amd_hybrid_plan and gc_amd_hybrid_update; nothing in
the product imports it.

The reference keeps these steps in its caller, batch mode 3 of
src/geos-chem/GeosCore/fullchem_mod.F90.  Each function below follows the
Fortran loop by loop (1-based slots and cells, as there); the tests pin this
restatement with checks that do not retype it (numpy's stable argsort,
brute-force prefix searches, hand-computed cases).
"""
from __future__ import annotations


def cost_slots(prev_nsteps, sort=True):
    """Slot of every cell (1-based) and the cell in every slot, by a stable
    counting sort on the previous step count, most expensive first
    (fullchem_mod.F90:917-963: pass 1 counts keys sMax - prev, the prefix sums
    give each key's first slot, pass 2 assigns slots in stock order)."""
    n = len(prev_nsteps)
    s_max = max(prev_nsteps) if (sort and n) else 0
    count = [0] * (s_max + 2)
    keys = []
    for c in range(n):                                   # :924-943
        key = s_max - prev_nsteps[c] if sort else 0
        count[key] += 1
        keys.append(key)
    pos = 1                                              # :945-950
    for key in range(s_max + 1):
        idx = count[key]
        count[key] = pos
        pos += idx
    slot_of = [0] * n
    cell_at = [0] * n
    for c in range(n):                                   # :952-962
        slot_of[c] = count[keys[c]]
        count[keys[c]] += 1
        cell_at[slot_of[c] - 1] = c + 1
    return slot_of, cell_at


def hybrid_plan(slot_nsteps, work_frac, rate_gpu, rate_cpu, ndev, dev_rate=None,
                fixed_frac=False):
    """(work, work_frac, nGpu, nDevUse, devLo, devHi) of fullchem_mod.F90:997-1061."""
    n = len(slot_nsteps)
    work = [0] * (n + 1)                                 # :999-1005
    for ic in range(1, n + 1):
        work[ic] = work[ic - 1] + max(slot_nsteps[ic - 1], 1)
    w_tot = float(work[n])
    if not fixed_frac and rate_gpu > 0.0 and rate_cpu > 0.0:   # :1007-1011
        work_frac = rate_gpu / (rate_gpu + rate_cpu)
        work_frac = min(max(work_frac, 0.01), 0.99)
    work_frac = min(max(work_frac, 0.0), 1.0)            # :1012
    w_target = work_frac * w_tot
    n_gpu = 0
    for ic in range(1, n + 1):                           # :1014-1017
        if float(work[ic]) > w_target:
            break
        n_gpu = ic
    n_dev_use = ndev                                     # :1021-1022
    if n_gpu < n_dev_use:
        n_dev_use = max(n_gpu, 1)
    dev_lo = [1] * ndev
    dev_hi = [0] * ndev
    dev_w = [1.0] * ndev                                 # :1032-1035
    if dev_rate is not None and all(dev_rate[d] > 0.0 for d in range(n_dev_use)):
        dev_w[:n_dev_use] = list(dev_rate[:n_dev_use])
    dev_sum = 0.0
    for d in range(n_dev_use):
        dev_sum += dev_w[d]
    idx = 0                                              # :1039-1057
    dev_cum = 0.0
    for d in range(n_dev_use):
        dev_lo[d] = idx + 1
        if d == n_dev_use - 1:
            dev_hi[d] = n_gpu
        else:
            dev_cum += dev_w[d]
            w_target = float(work[n_gpu]) * dev_cum / dev_sum
            for ic in range(idx + 1, n_gpu + 1):
                if float(work[ic]) > w_target:
                    break
                idx = ic
            dev_hi[d] = idx
        idx = dev_hi[d]
    return work, work_frac, n_gpu, n_dev_use, dev_lo, dev_hi


def hybrid_update(n_cell, n_gpu, work, n_dev_use, dev_lo, dev_hi, dev_tw, gpu_tw, t_cpu,
                  rate_gpu, rate_cpu, dev_rate):
    """Rates after a split solve: exponential moving averages of
    fullchem_mod.F90:2672-2700.  Returns (rate_gpu, rate_cpu, dev_rate)."""
    dev_rate = list(dev_rate)
    for d in range(n_dev_use):                           # :2673-2683
        if dev_hi[d] >= dev_lo[d] and dev_tw[d] > 0.0:
            r_new = float(work[dev_hi[d]] - work[dev_lo[d] - 1]) / dev_tw[d]
            dev_rate[d] = 0.5 * dev_rate[d] + 0.5 * r_new if dev_rate[d] > 0.0 else r_new
    if n_gpu > 0 and gpu_tw > 0.0:                       # :2684-2691
        r_new = float(work[n_gpu]) / gpu_tw
        rate_gpu = 0.5 * rate_gpu + 0.5 * r_new if rate_gpu > 0.0 else r_new
    if n_cell - n_gpu > 0 and t_cpu > 0.0:               # :2692-2699
        r_new = (float(work[n_cell]) - float(work[n_gpu])) / t_cpu
        rate_cpu = 0.5 * rate_cpu + 0.5 * r_new if rate_cpu > 0.0 else r_new
    return rate_gpu, rate_cpu, dev_rate
