"""MI355X (gfx950) mutation placement and sequencing simulation behind the
ProCESS R-level API. The hot path runs in libpsx.so (HIP kernels + C ABI,
include/psx.h); there is no CPU fallback."""
from .api import (BasicIlluminaSequencer, ErrorlessIlluminaSequencer, MutationEngine, PhylogeneticForest,
                  ReferenceGenome, SampledCell, SamplesForest, SNV, get_random_seed, get_seq_data, seq_baf, seq_dr,
                  seq_to_long, simulate_normal_seq, simulate_seq, validate_chromosomes)

__all__ = ["BasicIlluminaSequencer", "ErrorlessIlluminaSequencer", "MutationEngine", "PhylogeneticForest",
           "ReferenceGenome", "SampledCell", "SamplesForest", "SNV", "get_random_seed", "get_seq_data", "seq_baf",
           "seq_dr", "seq_to_long", "simulate_normal_seq", "simulate_seq", "validate_chromosomes"]
