"""mnpbem_amd: MI355X-native retarded BEM solve path of MNPBEM.

Drop-in names for the reference path ComParticle -> CompGreenRet -> BEMRet.solve ->
PlaneWaveRet.extinction. Host setup (meshes, quadrature, near-field coefficients) is numpy;
every per-wavelength numeric stage runs in libmnpbem_hip.so on the GPU.
"""
from .materials import EpsConst, EpsDrude, EpsTable
from .particle import Particle
from .mesh import trisphere, trirod, trispheresegment, triellipsoid, fvgrid
from .comparticle import ComParticle
from .compoint import ComPoint, MeshField
from .greenfun import CompGreenRet, CompGreenStat, CompStruct
from .bem import BEMRet
from .bem_stat import BEMStat, PlaneWaveStat
from .bem_stat_iter import ACACompGreenStat, BEMStatIter
from .bem_iter import BEMRetIter
from .planewave import PlaneWaveRet, SpectrumRet

__all__ = ["EpsConst", "EpsDrude", "EpsTable", "Particle", "trisphere", "trirod", "trispheresegment",
           "triellipsoid", "fvgrid", "ComParticle", "ComPoint", "MeshField", "CompGreenRet", "CompGreenStat", "CompStruct", "BEMRet", "BEMRetIter", "BEMStat",
           "BEMStatIter", "ACACompGreenStat", "PlaneWaveStat", "PlaneWaveRet",
           "SpectrumRet"]
