"""Many independent small systems evaluated together.

The reference evaluates one System per OpenMM Context; a fitting workload
that scores many small periodic systems (a few hundred to a few thousand
atoms each) creates one Context per system and calls getState on each in
turn (ReferenceCalcMPIDForceKernel::execute,
MPIDReferenceKernels.cpp:225-238, per call).  On one MI355X a single
small system fills a few dozen of the 256 CUs and its solver is bound by
kernel-launch latency, so one system after another leaves the chip idle.

``BatchEvaluator`` keeps one ``HipCalcMPIDForceKernel`` (its own three
streams, lists, grids and solver graph) per system and drives them from a
pool of host threads: the C ABI releases the GIL for the whole evaluation,
so the launches of several systems are in flight at once and their kernels
share the chip.  Each context issues its work on a single HIP stream: the
contexts share the process's few hardware queues, and a context's
cross-stream waits would hold up the other contexts' kernels queued behind
them.  Systems are handed out largest first (longest-processing-
time order) so that the batch does not end on one big system.  Every result
agrees with the one the system's own context gives to the rounding of the
grid atomics (same kernels, same fixed-point accumulation; the untiled PME
spread adds into the grid with floating-point atomics whose order varies),
whatever the thread count or order.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .kernel import HipCalcMPIDForceKernel
from .mpidforce import MPIDForce


class BatchEvaluator:
    """Energies, forces and induced dipoles of a list of independent systems.

    systems: a list of (MPIDForce, positions [n,3] nm, box) as the named
    workload generators return them.  threads: host threads issuing work
    (each blocks while its system's solver runs).  single_stream: each
    context issues its work on one stream (see MPID_FLAG_SINGLE_STREAM).
    """

    def __init__(self, systems, precision="mixed", threads=16, conservative_torques=False, single_stream=True):
        if threads < 1:
            raise ValueError("threads must be at least 1")
        self.precision = precision
        self.threads = int(threads)
        self.kernels = []
        self.positions = []
        self.boxes = []
        for force, pos, box in systems:
            if not isinstance(force, MPIDForce):
                raise TypeError("each system must be (MPIDForce, positions, box)")
            k = HipCalcMPIDForceKernel(precision, conservative_torques, single_stream)
            k.initialize(force.getNumMultipoles(), force, default_box=box)
            self.kernels.append(k)
            self.positions.append(np.ascontiguousarray(pos, dtype=np.float64))
            self.boxes.append(box)
        sizes = np.array([k._n for k in self.kernels])
        self.order = [int(i) for i in np.argsort(-sizes, kind="stable")]
        self._pool = ThreadPoolExecutor(max_workers=self.threads) if self.threads > 1 else None

    def __len__(self):
        return len(self.kernels)

    @property
    def num_atoms(self):
        return int(sum(k._n for k in self.kernels))

    def set_positions(self, i, positions, box=None):
        self.positions[i] = np.ascontiguousarray(positions, dtype=np.float64)
        if box is not None:
            self.boxes[i] = box

    def _map(self, fn, sequential=False):
        out = [None] * len(self.kernels)

        def run(i):
            out[i] = fn(i)

        if self._pool is None or sequential:
            for i in self.order:
                run(i)
        else:
            # consume the iterator so that the first error is raised here
            for _ in self._pool.map(run, self.order):
                pass
        return out

    def execute(self, include_forces=True, include_energy=True, sequential=False):
        """[(energy kJ/mol, forces [n,3] kJ/mol/nm)] in system order.
        sequential: one system after another on the calling thread."""
        return self._map(lambda i: self.kernels[i].execute(self.positions[i], self.boxes[i], include_forces,
                                                           include_energy), sequential)

    def energies(self):
        return np.array([e for e, _ in self.execute(include_forces=False)])

    def induced_dipoles(self):
        """[induced dipoles [n,3] e nm] in system order."""
        return self._map(lambda i: self.kernels[i].getInducedDipoles(self.positions[i], self.boxes[i]))

    def close(self):
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None
        self.kernels = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
