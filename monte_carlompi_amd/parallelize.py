"""``Parallel`` with the interface of monte_carloMPI/parallelize.py:6-41.

The reference splits the photon wavelengths over MPI ranks
(``np.array_split`` + ``comm.scatter``, parallelize.py:14-15, 28-38) and
gathers the per-photon answers back (``comm.gather``, :17-26).  Here one
process drives one GPU, which walks every photon of the working set in one
launch, so the communicator has a single rank: ``working_set`` is the whole
data set and ``answer_and_reduce`` applies the reduce function to the single
rank's answer.  ``split`` keeps the reference's array_split order for code
that shards photons over several processes.
"""
import numpy as np


class Parallel(object):
    def __init__(self, data_set, size=1, rank=0):
        self.size = size
        self.rank = rank
        self.working_set = self._map(data_set)

    def _partition(self, data_set, partitions):
        return np.array_split(data_set, partitions)

    def _map(self, data_set):
        return self._partition(data_set, self.size)[self.rank]

    def _reduce(self, answer, reduce_work_fn):
        return reduce_work_fn([answer]) if self.rank == 0 else None

    def answer_and_reduce(self, answer, reduce_work_fn):
        return self._reduce(answer, reduce_work_fn)
