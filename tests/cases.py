"""Seeded synthetic inputs shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

from process_amd import core


def signatures(n: int, seed: int) -> np.ndarray:
    """n SBS signatures, each a Dirichlet(1) draw over the 96 classes."""
    return np.random.default_rng(seed).dirichlet(np.ones(96), size=n)


def forest(n_cells=400, n_samples=2, sample_size=80, seed=3, two_species=True):
    if two_species:
        return core.synth_forest(growth=[0.2, 0.3], death=[0.01, 0.01], intro_at=[0, n_cells // 4],
                                 n_cells=n_cells, n_samples=n_samples, sample_size=sample_size, seed=seed)
    return core.synth_forest(growth=[0.2], death=[0.01], intro_at=[0], n_cells=n_cells, n_samples=n_samples,
                             sample_size=sample_size, seed=seed)


def base_of(packed: np.ndarray, lengths, chr_: int, pos: int) -> int:
    off = int(sum((int(l) + 15) // 16 for l in lengths[:chr_]))
    return int((packed[off + pos // 16] >> (2 * (pos % 16))) & 3)


def engine(packed, lengths, *, snv=2e-6, indel=2e-7, n_sig=2, epochs=(0.0, 5.0), pre_snv=4, pre_indel=2,
           infinite_sites=True, drivers=True, germline=True, sig_seed=1):
    """Engine parameters: two species, timed exposures over n_sig signatures."""
    sig = signatures(n_sig, sig_seed)
    ne = len(epochs)
    expo = np.random.default_rng(sig_seed + 1).dirichlet(np.ones(n_sig), size=ne)
    drv = []
    if drivers:
        pos = int(lengths[0]) // 2
        drv.append((1, 0, pos, (base_of(packed, lengths, 0, pos) + 1) % 4, -1))
    grm = []
    if germline:
        for k, pos in enumerate((101, int(lengths[-1]) - 50)):
            c = 0 if k == 0 else len(lengths) - 1
            grm.append((c, pos, (base_of(packed, lengths, c, pos) + 2) % 4, k % 2))
    return dict(snv_rate=[snv, snv * 1.5], indel_rate=[indel, indel], sbs=sig, epoch_start=list(epochs),
                exposure=expo, n_pre_snv=pre_snv, n_pre_indel=pre_indel, pre_sbs=0, indel_max_len=10,
                indel_len_p=0.5, infinite_sites=infinite_sites, ploidy=2, drivers=drv, germline=grm)


# Sequencing configurations exercised by the parity tests.
SEQ_CASES = {
    "se_errorless": dict(coverage=30, read_size=100),
    "pe_errors": dict(coverage=40, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3),
    "pe_high_error": dict(coverage=25, read_size=120, insert_size_mean=300, insert_size_stddev=12, error_rate=0.05),
    "purity": dict(coverage=30, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3,
                   purity=0.6, preneoplastic_in_normal=True),
    "subset": dict(coverage=20, read_size=80, error_rate=2e-3, chromosomes=[1], samples=[1],
                   with_normal_sample=False),
    "short_reads": dict(coverage=15, read_size=7, error_rate=0.01),
}
