"""qdfn — MI355X-native (gfx950) earthquake-cycle hot path of Quake-DFN.

Host-side mirror of the reference's Julia interface over the C ABI of
libqdfn_hip.so (include/qdfn.h).  The compute runs in hand-written HIP kernels;
this package only marshals arrays.  Importing it does not load the library;
the first call does, and raises if the library has not been built.
"""
from . import _lib
from ._lib import PREC_BF16_FAR, PREC_FP32, PREC_FP64, PREC_MIXED, PREC_MIXED8, QdfnError, device_info, lib
from .assembly import (BuildMatrixByParts_Even_Rec, BuildMatrixByParts_Even_Tri, Okada_DC3D_Vector,
                       ReduceTooStrongInteraction, RotVerts_UnitVectors, StiffnessMatrix_ByParts_Calculation_Rec,
                       StiffnessMatrix_ByParts_Calculation_Tri, TDstressHS)
from .catalog import EventCatalog, fault_area, get_event_fragments, merge_fragments
from .ensemble import Ensemble
from .hmat import HMatrix, ReduceTooStrongInteraction_Hmat, hmat_structure, main_H
from .inputs import BulkToSegment, LRtoRake, ReadBulkInput, ReadSegmentInput
from .sim import DenseOperator, Simulation, main
from .solvers import SolveOneTimeStep

__all__ = ["PREC_BF16_FAR", "PREC_FP32", "PREC_FP64", "PREC_MIXED", "PREC_MIXED8", "QdfnError", "device_info", "lib", "Okada_DC3D_Vector",
           "StiffnessMatrix_ByParts_Calculation_Rec", "BuildMatrixByParts_Even_Rec", "TDstressHS",
           "RotVerts_UnitVectors", "StiffnessMatrix_ByParts_Calculation_Tri", "BuildMatrixByParts_Even_Tri",
           "DenseOperator", "Simulation", "main", "SolveOneTimeStep", "HMatrix", "hmat_structure", "main_H",
           "BulkToSegment", "LRtoRake", "ReadBulkInput", "ReadSegmentInput",
           "EventCatalog", "fault_area", "get_event_fragments", "merge_fragments", "Ensemble",
           "ReduceTooStrongInteraction", "ReduceTooStrongInteraction_Hmat"]
