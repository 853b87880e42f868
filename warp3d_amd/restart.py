"""Element-block state on disk in the reference's restart records.

The reference saves the element state of every block in the restart file
(store.f:929-997, read back in the same order by reopen.f) as Fortran
unformatted sequential records: each block's array, in its in-memory layout,
followed by a record holding the check key 2147483647.  The sections, in
the order store.f writes them:

  1  history_blocks (hist, ngp, span) of the blocks with history, each
     followed by the block's cep_blocks (21, span, ngp);
  2  rot_n1_blocks (9, ngp, span) of the finite-strain blocks;
  4  eps_n_blocks (6, ngp, span) of every block;
  5  urcs_n_blocks (9, ngp, span) of every block.

(Section 3, the initial-state arrays, belongs to a feature that is out of
scope here.)  A record is a 4-byte little-endian length, the data and the
length again (gfortran's and ifort's default unformatted layout).  The
arrays are those of a committed state, as store.f sees them after update.

Both the device model (warp3d_amd.model.BlockModel) and the oracle use the
same block interface (nelblk, block_info, get_block_state,
set_block_state), so a state written from one can be read into the other.
"""
import struct

import numpy as np

CHECK_KEY = 2147483647

# selectors of w3d_get_block_state
HIST_N, CEP, STRAIN_N, STRESS_N, ROT_N1 = 4, 6, 2, 0, 10


class RestartFormatError(ValueError):
    pass


def write_record(f, data):
    """One unformatted sequential record of raw bytes."""
    b = bytes(data)
    n = struct.pack("<i", len(b))
    f.write(n)
    f.write(b)
    f.write(n)


def read_record(f):
    head = f.read(4)
    if len(head) < 4:
        raise RestartFormatError("end of file inside the element sections")
    (n,) = struct.unpack("<i", head)
    b = f.read(n)
    tail = f.read(4)
    if len(b) != n or tail != head:
        raise RestartFormatError("broken record length markers")
    return b


def _key(f):
    write_record(f, struct.pack("<i", CHECK_KEY))


def _check_key(f):
    b = read_record(f)
    if len(b) != 4 or struct.unpack("<i", b)[0] != CHECK_KEY:
        raise RestartFormatError("missing check key after a block record")


def _blocks(model, which):
    """(block, selector) pairs of the records of section 1, 4 or 5, in
    store.f's order."""
    out = []
    for b in range(1, model.nelblk + 1):
        if which == 1:
            if model.block_info(b)["hist"] > 0:
                out.append((b, HIST_N))
            out.append((b, CEP))
        elif which == 4:
            out.append((b, STRAIN_N))
        elif which == 5:
            out.append((b, STRESS_N))
    return out


def save_element_state(model, path, sections=(1, 2, 4, 5), geonl=None):
    """Write the element sections of a committed state.  geonl: per block
    (1-based list index 0 = block 1) whether it has finite strains; the
    rotation section is written for those blocks."""
    with open(path, "wb") as f:
        for sec in sections:
            for b, which in _records(model, sec, geonl):
                a = np.ascontiguousarray(model.get_block_state(b, which),
                                         dtype="<f8")
                write_record(f, a.tobytes())
                _key(f)


def _records(model, sec, geonl):
    if sec == 2:
        if geonl is None:
            return []
        return [(b, ROT_N1) for b in range(1, model.nelblk + 1)
                if geonl[b - 1]]
    return _blocks(model, sec)


def load_element_state(model, path, sections=(1, 2, 4, 5), geonl=None):
    """Read the element sections written for this model's blocking.  The
    n-state arrays (history, strains, stresses) are set on the model, and
    copied to the n+1 state as after a commit; the
    cep and rotation arrays, which the next recovery recomputes, are
    returned as {(block, selector): array}.  Every record's size must be
    the model's block size."""
    rest = {}
    with open(path, "rb") as f:
        for sec in sections:
            for b, which in _records(model, sec, geonl):
                raw = read_record(f)
                n = model.state_size(b, which)
                if len(raw) != 8 * n:
                    raise RestartFormatError(
                        f"block {b}: record of {len(raw)} bytes, the model's "
                        f"array has {8 * n}")
                a = np.frombuffer(raw, dtype="<f8").astype(np.float64)
                _check_key(f)
                if which in (HIST_N, STRAIN_N, STRESS_N):
                    model.set_block_state(b, which, a)
                else:
                    rest[(b, which)] = a
        if f.read(1):
            raise RestartFormatError("data after the element sections")
    # a committed state has n+1 = n (update.f copies one into the other);
    # the recovery of the next step reads both
    model.rollback()
    return rest
