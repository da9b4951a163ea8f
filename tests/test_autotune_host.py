"""The autotuner's sample of a batch (biomedisa_amd/autotune.py): the leading
volumes holding about 1/nbrw of the seeds; no device needed."""
import numpy as np

from biomedisa_amd import _hip
from biomedisa_amd.autotune import sample_table


class FakeBatch:
    def __init__(self, counts):
        self.nseeds = list(counts)
        self.table = np.zeros(len(counts), dtype=_hip.VOLUME_DTYPE)


def test_leading_volumes_with_a_tenth_of_the_seeds():
    b = FakeBatch([100] * 50)
    t = sample_table(b, 10)
    assert len(t) == 5                                      # the first volumes reaching 1/10 of the seeds


def test_uneven_volumes():
    t = sample_table(FakeBatch([10, 60, 10] + [100] * 7), 10)
    assert len(t) == 3                                      # 70 < 78 <= 80
    # reaching 1/10 would take a volume of most of the seeds: no sample
    assert sample_table(FakeBatch([10, 10, 10, 500, 10, 10]), 10) is None


def test_no_proper_subset():
    assert sample_table(FakeBatch([1000]), 10) is None      # one volume
    assert sample_table(FakeBatch([900, 10, 10]), 10) is None  # the first volume alone is most of the work
    assert sample_table(FakeBatch([5, 5]), 1) is None       # nbrw 1: the whole walk
    assert sample_table(FakeBatch([0, 0]), 10) is None
