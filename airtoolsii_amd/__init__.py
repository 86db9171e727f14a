"""airtoolsii_amd -- an MI355X-native engine for the SIRT methods of AIR Tools II.

Call surface (AIR Tools II, MATLAB -> Python):
    sirt, landweber, cimmino, cav, drop, sart          sirt/*.m
    paralleltomo, fanlineartomo, fancurvedtomo,
    seismictomo, sphericaltomo, seismicwavetomo        testprobs/*.m (built on the GPU)
    phantomgallery                                     testprobs/phantomgallery.m
    train_relaxpar, train_dpme                         train/train_relaxpar.m, train/train_dpme.m
    fbp                                                auxil/fbp.m
    kaczmarz, symkaczmarz, randkaczmarz, art           art/*.m (one wave per sweep)
    columnaction, cart                                 cart/*.m
Operators follow the afun convention A(v,'notransp'|'transp'|'size').
The compute runs in libairhip.so (HIP kernels for gfx950, C ABI in
include/airhip.h); build it with ``python -m airtoolsii_amd.build``.
"""
from .batch import BatchOperator, batch_sirt, paralleltomo_batch
from .operators import AirOperator, FunctionOperator, as_operator, from_matlab_csc
from .phantoms import add_noise, phantomgallery
from .sirt import (SirtRun, calczeta, cav, cimmino, drop, landweber, sart, sirt, sirt_multi, spectral_radius,
                   weights)
from .testprobs import (fancurvedtomo, fanlineartomo, paralleltomo, seismictomo, seismicwavetomo,
                        sphericaltomo)
from .art import art, cart, columnaction, kaczmarz, randkaczmarz, symkaczmarz
from .fbp import fbp
from .train import train_dpme, train_relaxpar

__all__ = ["BatchOperator", "batch_sirt", "paralleltomo_batch", "AirOperator", "FunctionOperator", "as_operator", "from_matlab_csc", "add_noise", "phantomgallery", "SirtRun",
           "calczeta", "cav", "cimmino", "drop", "landweber", "sart", "sirt", "sirt_multi", "spectral_radius", "weights",
           "fancurvedtomo", "fanlineartomo", "paralleltomo", "seismictomo", "seismicwavetomo", "sphericaltomo", "train_relaxpar", "train_dpme", "fbp", "art", "cart", "columnaction",
           "kaczmarz", "randkaczmarz", "symkaczmarz"]
