"""MI355X (gfx950) implementation of the TDFstat coherent F-statistic search
core (search/network job_core path).  The compute path is libtdfstat_hip.so
(C ABI in include/tdfstat_hip.h); this package holds its Python binding, the
host-side setup (settings, synthetic inputs) and the job_core/search driver.
"""
__all__ = ["settings", "synth", "search"]
