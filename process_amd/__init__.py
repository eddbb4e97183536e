"""MI355X (gfx950) mutation placement and sequencing simulation behind the
ProCESS R-level API. The hot path runs in libpsx.so (HIP kernels + C ABI,
include/psx.h); there is no CPU fallback."""
from .api import (CNA, WGD, Amplification, BasicIlluminaSequencer, Deletion, ErrorlessIlluminaSequencer,
                  MutationEngine, PhylogeneticForest, ReferenceGenome, SampledCell, SamplesForest, SNV,
                  get_random_seed, get_seq_data, load_passenger_CNAs, load_phylogenetic_forest, seq_baf, seq_dr, seq_to_long,
                  simulate_normal_seq, simulate_seq, validate_chromosomes)

__all__ = ["Amplification", "BasicIlluminaSequencer", "CNA", "Deletion", "ErrorlessIlluminaSequencer",
           "MutationEngine", "PhylogeneticForest", "ReferenceGenome", "SampledCell", "SamplesForest", "SNV", "WGD",
           "get_random_seed", "get_seq_data", "load_passenger_CNAs", "load_phylogenetic_forest", "seq_baf", "seq_dr", "seq_to_long",
           "simulate_normal_seq", "simulate_seq", "validate_chromosomes"]
