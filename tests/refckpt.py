"""Writer of the reference's checkpoint format, for tests: the layout of
writeCheckpoint (gpuLucas.cu:1641-1665) with the weighted digits x_j * a_j
that the reference keeps as its next FFT input."""
import numpy as np

from gpulucas_amd import lucas


def write_reference_checkpoint(path, p, N, loop_index, digits, name=b"gpuLucas", version=b"0.9.4"):
    """State s_(loop_index - 1) given as its digits, as the reference writes it
    after loop index `loop_index`."""
    _, a, _ = lucas.digit_tables(p, N)
    signal = np.asarray(digits, dtype=np.float64) * a
    with open(path, "wb") as f:
        f.write(name.ljust(16, b"\0"))
        f.write(version.ljust(16, b"\0"))
        f.write(np.array([p, N, loop_index], dtype=np.int32).tobytes())
        f.write(signal.astype("<f8").tobytes())
