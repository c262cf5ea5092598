"""FP64 CPU oracle for the multislice hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in double precision numpy/scipy, what the reference
nlaanait/namsa MSAGPU path computes (namsa/msa.py, namsa/probe_kernels.cu,
namsa/potential_kernels.cu, namsa/optics.py, namsa/utils.py).  It is the
checker for the HIP product path, never part of it: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package ``namsa_amd`` never imports anything from here.

Modules
-------
msa_oracle : MSAGPU conventions (slice builder, probe, propagator, multislice,
             readout, detectors, probe raster) in FP64.
cpu_port   : op-for-op restatement of the reference CPU ``MSA.propagate_beam``
             numpy path (complex64, descending slice order); used only as the
             timed CPU baseline and checked against ``msa_oracle`` where the
             two conventions coincide.

Pinning status of every function (which test pins it, and how) is listed in
the header of each module.
"""
