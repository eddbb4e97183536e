"""The C ABI on the CPU: libpsx loads and exports every entry point declared
in include/psx.h, struct layouts agree with the header, and the host-only
logic (result table, insert-size validation, synthetic forests) matches the
oracle and the reference's messages."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from process_amd import _abi as A
from process_amd import core
from process_amd._lib import LIB_PATH, PsxError, check, lib

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "psx.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(psx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (psx_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    assert len(declared_functions()) >= 40
    assert lib().psx_abi_version() == 1


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "psx.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(psx_mutation), sizeof(psx_forest_desc),
         sizeof(psx_engine_params), sizeof(psx_seq_params), sizeof(psx_seq_job), sizeof(psx_launch_config),
         sizeof(psx_synth_params), sizeof(psx_driver), sizeof(psx_germline));
  printf("%zu %zu %zu %zu\\n", offsetof(psx_mutation, cause), offsetof(psx_mutation, sid_head),
         offsetof(psx_seq_params, chromosomes), offsetof(psx_engine_params, germline));
  printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(psx_server_config), sizeof(psx_request), sizeof(psx_request_stats),
         sizeof(psx_server_totals), offsetof(psx_request, seed), offsetof(psx_request_stats, latency_ms));
  return 0;
}""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = list(map(int, lines[0].split()))
    assert sizes == [A.MUTATION_DTYPE.itemsize, C.sizeof(A.ForestDesc), C.sizeof(A.EngineParams),
                     C.sizeof(A.SeqParams), C.sizeof(A.SeqJob), C.sizeof(A.LaunchConfig), C.sizeof(A.SynthParams),
                     C.sizeof(A.Driver), C.sizeof(A.Germline)]
    srv = list(map(int, lines[2].split()))
    assert srv == [C.sizeof(A.ServerConfig), C.sizeof(A.Request), C.sizeof(A.RequestStats),
                   C.sizeof(A.ServerTotals), A.Request.seed.offset, A.RequestStats.latency_ms.offset]
    offs = list(map(int, lines[1].split()))
    assert offs == [A.MUTATION_DTYPE.fields["cause"][1], A.MUTATION_DTYPE.fields["sid_head"][1],
                    A.SeqParams.chromosomes.offset, A.EngineParams.germline.offset]


def test_insert_size_error_message_matches_reference():
    """src/seq_simulation.cpp:437-445."""
    st = lib().psx_insert_size_check(None, 50, 10)
    with pytest.raises(PsxError, match=r"The insert size mean \(50\) must be greater than or equal to its "
                                       r"variance \(10\*10=100\)"):
        check(st)
    check(lib().psx_insert_size_check(None, 350, 10))


def test_context_without_gpu_fails_loudly():
    try:
        import torch
        has_gpu = torch.cuda.device_count() > 0
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(PsxError):
        core.Context(0)


def test_synthetic_forest_properties():
    kw = dict(growth=[0.2, 0.3], death=[0.01, 0.01], intro_at=[0, 1000], n_cells=5000, n_samples=3,
              sample_size=500)
    a = core.synth_forest(seed=5, **kw)
    b = core.synth_forest(seed=5, **kw)
    c = core.synth_forest(seed=6, **kw)
    for k in ("cell_id", "parent", "species", "sample", "birth"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["parent"], c["parent"])
    assert np.bincount(a["sample"][a["sample"] >= 0]).tolist() == [500, 500, 500]
    f = O.Forest(a["cell_id"], a["parent"], a["species"], a["sample"], a["birth"], 3)
    assert f.n_leaves == 1500  # the sampled cells are exactly the leaves
    # births are not earlier than the parent's
    by = dict(zip(a["cell_id"].tolist(), a["birth"].tolist()))
    for cid, par, t in zip(a["cell_id"], a["parent"], a["birth"]):
        if par >= 0:
            assert t >= by[int(par)]
    assert (a["species"] == 1).any()


def test_synthetic_forest_rejects_bad_parameters():
    with pytest.raises(PsxError, match="samples exceed"):
        core.synth_forest(growth=[0.2], death=[0.0], intro_at=[0], n_cells=100, n_samples=2, sample_size=60, seed=1)


@pytest.mark.parametrize("include", [False, True])
def test_result_table_matches_oracle(include):
    rng = np.random.default_rng(3)
    M, S = 400, 3
    m = np.zeros(M, A.MUTATION_DTYPE)
    m["pos"] = np.sort(rng.integers(0, 5000, M))
    m["alt"] = rng.integers(0, 4, M)
    m["kind"] = rng.choice([0, 0, 0, 1, 2], M)
    m = O.sort_table(m)
    alt = rng.integers(0, 3, (S, M)).astype(np.int32) * (rng.random((S, M)) < 0.3)
    dep = (alt + rng.integers(0, 30, (S, M))).astype(np.int32)
    dep[0, :5] = 0
    alt[0, :5] = 0
    pres = (rng.random((S, M)) < 0.4).astype(np.uint8)
    pres[0, :5] = 1
    ref = O.seq_table(m, alt, dep, pres, include)
    got = core.seq_table(m, alt, dep, pres, include)
    assert np.array_equal(ref[0], got[0])
    for x, y in zip(ref[1:], got[1:]):
        assert np.array_equal(x, y, equal_nan=True) if x.dtype.kind == "f" else np.array_equal(x, y)
    # VAF = occ/cov when held (NaN on zero coverage), else 0 (src/seq_simulation.cpp:123-130)
    if include:
        assert np.isnan(got[3]).any()
