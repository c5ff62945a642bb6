"""ymc3d_amd - photon-packet Monte Carlo in 3-D voxel tissue for AMD Instinct MI355X (gfx950).

The hot path (packet launch, voxel DDA, absorption, Henyey-Greenstein scattering, Fresnel at index
mismatches, roulette and the diffuse-reflectance tally) is a persistent HIP kernel in
libymc3d.so, reached through the C ABI of include/ymc3d.h.  This package is the host-side mirror
of the reference's driver: file formats, scene generators and the engine wrapper.
"""
from ._abi import FX_ONE, FRESNEL_EXACT, FRESNEL_YMC3D, ILLUM_PENCIL, ILLUM_UNIFORM, FLAG_COUNTERS, FLAG_RECORD, FLAG_ABSORPTION
from .engine import (Engine, TallyResult, YmcError, build_chains, default_multipliers, generate_multipliers,
                     layered_voxelize, load_library, read_multipliers_text, run_3dmc_gpu, run_once, seed_streams)

__all__ = ["Engine", "TallyResult", "YmcError", "build_chains", "default_multipliers", "generate_multipliers",
           "layered_voxelize", "load_library", "read_multipliers_text", "run_3dmc_gpu", "run_once", "seed_streams", "FX_ONE", "FRESNEL_EXACT",
           "FRESNEL_YMC3D", "ILLUM_PENCIL", "ILLUM_UNIFORM", "FLAG_COUNTERS", "FLAG_RECORD", "FLAG_ABSORPTION"]
