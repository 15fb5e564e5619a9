"""Drop-in for ``sinv.algorithms.bcr_s.bcr_serial`` (bcr_serial.py:19-62)."""


def bcr_serial(A, blocksize):
    raise NotImplementedError("GPU BCR-serial is not built yet")
