"""Event catalog on the device (scripts/Functions_Plot.jl:430-579).

  get_event_fragments(ResultTime, ResultDisp, FaultCenter, FaultArea, ShearModulus, DispRateCrits)
        the fragment pass of get_event_fragments (:430-510) on history arrays
        (the reference loads them from Result.jld2 / Result_Input.jld2)
  merge_fragments(...)   merge_fragments (:513-579), same arguments and returns
  fault_area(...)        FaultArea as get_event_fragments computes it (:453-465)
  EventCatalog           the handle: fragments straight from a run's
                         device-resident histories, merge on the device

Fragments open where the recorded slip rate rises above DispRateCrits and sum
dD * FaultArea * ShearModulus while it stays above; they are sorted by time.
Merging joins a fragment to the event of the nearest earlier fragment within
MergeTimeCriteria and MergeDistanceCriteria; Mw = (log10 M0 - 9.05) / 1.5.
"""
import ctypes as C

import numpy as np

from ._lib import I64, check, check_handle, lib, ptr, ptr_i32


def fault_area(RorT, FaultLengthStrike=None, FaultLengthDip=None, P1=None, P2=None, P3=None):
    """Element areas: Heron's formula on the triangle edges for "T", else
    FaultLengthStrike .* FaultLengthDip (:453-465)."""
    if RorT == "T":
        P1, P2, P3 = (np.asarray(P, dtype=np.float64) for P in (P1, P2, P3))
        a = np.sqrt(np.sum((P1 - P2) ** 2, axis=1))
        b = np.sqrt(np.sum((P1 - P3) ** 2, axis=1))
        c = np.sqrt(np.sum((P2 - P3) ** 2, axis=1))
        s = (a + b + c) / 2
        return np.sqrt(s * (s - a) * (s - b) * (s - c))
    return np.asarray(FaultLengthStrike, dtype=np.float64) * np.asarray(FaultLengthDip, dtype=np.float64)


class EventCatalog:
    """Device-resident fragments and merged events."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_sim(cls, sim, FaultCenter, FaultArea, ShearModulus, DispRateCrits=1e-2):
        """Fragments from the run's own History_Time / History_Disp records,
        without copying the histories to the host."""
        ctr = np.asfortranarray(FaultCenter, dtype=np.float64)
        area = np.ascontiguousarray(FaultArea, dtype=np.float64)
        return cls(check_handle(lib().qdfn_catalog_from_sim(sim._h, ptr(ctr), ptr(area), float(ShearModulus),
                                                           float(DispRateCrits)), "qdfn_catalog_from_sim"))

    @classmethod
    def from_history(cls, ResultTime, ResultDisp, FaultCenter, FaultArea, ShearModulus, DispRateCrits=1e-2):
        t = np.ascontiguousarray(ResultTime, dtype=np.float64).ravel()
        d = np.asfortranarray(ResultDisp, dtype=np.float64)
        if d.ndim != 2 or d.shape[0] != t.size:
            raise ValueError("ResultDisp must be n_records x n with n_records = len(ResultTime)")
        n = d.shape[1]
        ctr = np.asfortranarray(FaultCenter, dtype=np.float64)
        area = np.ascontiguousarray(FaultArea, dtype=np.float64)
        if ctr.shape != (n, 3) or area.shape != (n,):
            raise ValueError("FaultCenter must be n x 3 and FaultArea length n")
        h = lib().qdfn_catalog_from_history(t.size, n, ptr(t), ptr(d), max(t.size, 1), ptr(ctr), ptr(area),
                                            float(ShearModulus), float(DispRateCrits))
        return cls(check_handle(h, "qdfn_catalog_from_history"))

    @classmethod
    def from_fragments(cls, EventTime_Fragment, EventLocation_Fragment, EventMoment_Fragment):
        t = np.ascontiguousarray(EventTime_Fragment, dtype=np.float64).ravel()
        F = t.size
        loc = np.asfortranarray(np.asarray(EventLocation_Fragment, dtype=np.float64).reshape(F, 3))
        m = np.ascontiguousarray(EventMoment_Fragment, dtype=np.float64).ravel()
        if m.size != F:
            raise ValueError("fragment arrays differ in length")
        return cls(check_handle(lib().qdfn_catalog_from_fragments(F, ptr(t), ptr(loc), ptr(m)),
                                "qdfn_catalog_from_fragments"))

    def fragments(self):
        """(EventTime_Fragment, EventLocation_Fragment (F x 3), EventMoment_Fragment,
        element index 0-based, -1 for fragments given directly)."""
        F = I64()
        check(lib().qdfn_catalog_fragments(self._h, C.byref(F), None, None, None, None))
        F = F.value
        t, loc, m = np.zeros(F), np.zeros((F, 3), order="F"), np.zeros(F)
        e = np.zeros(F, dtype=np.int32)
        if F:
            check(lib().qdfn_catalog_fragments(self._h, C.byref(I64()), ptr(t), ptr(loc), ptr(m), ptr_i32(e)))
        return t, np.ascontiguousarray(loc), m, e

    def merge(self, MergeTimeCriteria, MergeDistanceCriteria):
        E = I64()
        check(lib().qdfn_catalog_merge(self._h, float(MergeTimeCriteria), float(MergeDistanceCriteria),
                                       C.byref(E)))
        return E.value

    def events(self):
        """(EventTime_Merged, EventMomentMagnitude, EventLocation_Merged (E x 3),
        EventMoment_Bulk, event of each fragment 0-based)."""
        E = I64()
        check(lib().qdfn_catalog_events(self._h, C.byref(E), None, None, None, None, None))
        E = E.value
        F = I64()
        check(lib().qdfn_catalog_fragments(self._h, C.byref(F), None, None, None, None))
        t, mw, loc, m = np.zeros(E), np.zeros(E), np.zeros((E, 3), order="F"), np.zeros(E)
        fid = np.zeros(F.value, dtype=np.int32)
        if E:
            check(lib().qdfn_catalog_events(self._h, C.byref(I64()), ptr(t), ptr(mw), ptr(loc), ptr(m),
                                            ptr_i32(fid)))
        return t, mw, np.ascontiguousarray(loc), m, fid

    def close(self):
        if self._h:
            lib().qdfn_catalog_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def get_event_fragments(ResultTime, ResultDisp, FaultCenter, FaultArea, ShearModulus, DispRateCrits=1e-2):
    """get_event_fragments (:430-510) on arrays: returns (EventTime_Fragment,
    EventLocation_Fragment, EventMoment_Fragment, EventCount_Fragment)."""
    with EventCatalog.from_history(ResultTime, ResultDisp, FaultCenter, FaultArea, ShearModulus,
                                   DispRateCrits) as c:
        t, loc, m, _ = c.fragments()
    return t, loc, m, t.size


def merge_fragments(EventTime_Fragment, EventLocation_Fragment, EventMoment_Fragment, EventCount_Fragment,
                    MergeTimeCriteria, MergeDistanceCriteria):
    """merge_fragments (:513-579): returns (EventTime_Merged, EventMomentMagnitude,
    EventLocation_Merged).  With no fragments the reference returns its
    placeholders [0] and [0 0 0] with an empty magnitude list; so does this."""
    F = int(EventCount_Fragment)
    t = np.asarray(EventTime_Fragment, dtype=np.float64).ravel()[:F]
    loc = np.asarray(EventLocation_Fragment, dtype=np.float64).reshape(-1, 3)[:F]
    m = np.asarray(EventMoment_Fragment, dtype=np.float64).ravel()[:F]
    if F == 0:
        return np.zeros(1), np.zeros(0), np.zeros((1, 3))
    with EventCatalog.from_fragments(t, loc, m) as c:
        c.merge(MergeTimeCriteria, MergeDistanceCriteria)
        tm, mw, lm, _, _ = c.events()
    return tm, mw, lm
