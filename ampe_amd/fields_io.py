"""Initial fields from NetCDF files, with the semantics of the reference's
reader FieldsInitializer::initializeLevelFromData / initializePatchFromData
(source/FieldsInitializer.cc:77-665, NETCDF4 branch: the one the reference
compiles, source/FieldsInitializer.h:14), and a writer in the layout of the
reference's generator utils/make_nuclei.py (:438-520).

Format.  netcdf-cxx4 reads every NetCDF format; this image has no NetCDF or
HDF5 library, only scipy's reader and writer of the classic formats (CDF-1
and the 64-bit-offset CDF-2).  Classic files are read here; a NetCDF-4 file
(HDF5 container, what make_nuclei.py writes with format='NETCDF4') is refused
with a message naming the conversion (nccopy -k classic).  The writer makes
CDF-2 files with make_nuclei.py's dimension and variable names.

Restated behaviour, quirks kept:
  * order parameters: for each of the nphases orders the variable 'phase' is
    taken when it exists (every order then reads the same field), else
    'phase<i>' (FieldsInitializer.cc:150-163); when the state holds one order
    more than the file gives, the last one is 1 minus the sum of the others
    (:420-433);
  * 'temperature' is read only when asked for (readT, :166-172), else the
    uniform value fills the field (:527-534); quaternion components
    'quat1'..'quat<qlen>' are counted up to the first missing one (:173-181)
    and the count must equal qlen (:702-706); each must exist (:262-268);
  * composition: 'concentration' with one species, falling back to
    'concentration0'; 'concentration<i>' with several (:296-313);
  * the file's (z, y, x) extents come from the first phase variable (else
    the temperature) and must equal the domain's (nz = 1 in 2-D, :685-707);
    in 2-D the plane read is nz_file / 2 when slice_index < 0, else plane 0
    (:217-225: a non-negative slice index is not used);
  * the values are read in the precision of the first phase variable (or of
    the temperature): float files go through single precision, so a double
    variable in a float file is rounded to float (:320-334);
  * a patch reads the slab of its box (:372-392); reading the whole domain and
    cutting boxes from it gives the same values.
"""
import numpy as np

_HDF5_SIGNATURE = b"\x89HDF\r\n\x1a\n"


class FieldsFileError(RuntimeError):
    pass


def _open(path):
    from scipy.io import netcdf_file
    with open(path, "rb") as f:
        head = f.read(8)
    if head.startswith(_HDF5_SIGNATURE):
        raise FieldsFileError(
            f"{path}: NetCDF-4 (HDF5) file; this reader takes the classic formats only "
            "(convert with: nccopy -k classic in.nc out.nc)")
    if not head.startswith(b"CDF"):
        raise FieldsFileError(f"{path}: not a NetCDF file")
    return netcdf_file(path, "r", mmap=False)


def read_fields(path, domain, nphases=1, norder=None, qlen=4, ncompositions=1, read_q=True,
                read_c=True, read_t=False, read_eta=False, uniform_q=None, uniform_c=None,
                uniform_t=None, slice_index=-1):
    """Fields of the domain box from a NetCDF file.

    domain: the domain's cell counts, (nx, ny) in 2-D or (nx, ny, nz) in 3-D.
    nphases: order parameters in the file; norder: order parameters of the
    state (nphases or nphases + 1).  read_q / read_c / read_t: read the
    orientation / composition / temperature from the file, else fill them
    with uniform_q (qlen values) / uniform_c (ncompositions values) /
    uniform_t.  Returns a dict of float64 arrays shaped ([nz,] ny, nx) with a
    leading component axis for 'phase' (norder), 'quat' (qlen) and 'conc'
    (ncompositions), plus 'temperature' and 'eta' when present."""
    ndim = len(domain)
    if ndim not in (2, 3):
        raise ValueError("domain: (nx, ny) or (nx, ny, nz)")
    nx, ny = int(domain[0]), int(domain[1])
    nz = int(domain[2]) if ndim == 3 else 1
    norder = nphases if norder is None else norder
    if norder not in (nphases, nphases + 1):
        raise ValueError("norder must be nphases or nphases + 1")
    f = _open(path)
    try:
        V = f.variables

        def need(name):
            if name not in V:
                raise FieldsFileError(f"Could not read variable {name} from input data")
            return V[name]

        phase = []
        for i in range(nphases):
            if "phase" in V:
                phase.append(V["phase"])
            else:
                phase.append(need(f"phase{i}"))
        eta = need("eta") if read_eta else None
        temp = need("temperature") if read_t else None
        qlen_file = 0
        if read_q and qlen > 0:
            for m in range(qlen):
                if f"quat{m + 1}" not in V:
                    break
                qlen_file += 1
        # extents and precision from the first phase variable, else the
        # temperature
        ref = phase[0] if phase else temp
        if ref is None:
            nz_f, ny_f, nx_f = nz, ny, nx
            dt = np.float32
        else:
            if len(ref.dimensions) != 3:
                raise FieldsFileError("input variables must be (z, y, x)")
            nz_f, ny_f, nx_f = ref.shape
            dt = np.float64 if ref.data.dtype.itemsize == 8 else np.float32
        islice = 0
        if ndim == 2 and slice_index < 0:
            islice = nz_f // 2
        if (nx_f, ny_f, nz_f) != (nx, ny, nz):
            raise FieldsFileError(
                f"Phase input data dimensions are incorrect, nx_file={nx_f}, ny_file={ny_f}, "
                f"nz_file={nz_f}, nx_prob={nx}, ny_prob={ny}, nz_prob={nz}")
        if read_q and qlen > 0 and qlen_file != qlen:
            raise FieldsFileError(
                f"Phase input data dimensions are incorrect, qlen_file={qlen_file}, QLEN={qlen}")

        def slab(var):
            a = np.asarray(var.data)
            a = a[islice:islice + 1] if ndim == 2 else a
            # read in the reference's buffer precision, kept as double
            a = a.astype(dt).astype(np.float64)
            return a[0] if ndim == 2 else a

        shape = (ny, nx) if ndim == 2 else (nz, ny, nx)
        out = {}
        ph = np.zeros((norder,) + shape)
        for i in range(nphases):
            ph[i] = slab(phase[i])
        if norder > nphases:
            s = np.zeros(shape)
            for i in range(nphases):
                s = s + ph[i]
            ph[nphases] = 1.0 - s
        out["phase"] = ph
        if eta is not None:
            out["eta"] = slab(eta)
        if temp is not None:
            out["temperature"] = slab(temp)
        elif uniform_t is not None:
            out["temperature"] = np.full(shape, float(uniform_t))
        if qlen > 0:
            q = np.zeros((qlen,) + shape)
            if read_q:
                for m in range(qlen):
                    q[m] = slab(need(f"quat{m + 1}"))
            else:
                if uniform_q is None or len(uniform_q) != qlen:
                    raise FieldsFileError(f"need {qlen} values to specify q")
                for m in range(qlen):
                    q[m] = float(uniform_q[m])
            out["quat"] = q
        if ncompositions > 0:
            c = np.zeros((ncompositions,) + shape)
            for cc in range(ncompositions):
                if read_c:
                    name = "concentration" + (str(cc) if ncompositions > 1 else "")
                    if name not in V and ncompositions == 1:
                        name = name + "0"
                    c[cc] = slab(need(name))
                else:
                    if uniform_c is None or len(uniform_c) != ncompositions:
                        raise FieldsFileError(
                            f"Expect {ncompositions} to specify alloy concentration")
                    c[cc] = float(uniform_c[cc])
            out["conc"] = c
        return out
    finally:
        f.close()


def write_fields(path, phase, conc=None, quat=None, temperature=None, eta=None,
                 double_precision=False, phase_names=None):
    """A CDF-2 (64-bit offset) file in make_nuclei.py's layout: dimensions x,
    y, z, ns (species) and qlen; variables 'phase' (or phase_names, one per
    order), 'concentration<s>', 'quat<m>' (m from 1), 'temperature', 'eta',
    all (z, y, x), 'd' or 'f'.  Arrays are ([nz,] ny, nx); conc and quat
    carry a leading component axis (phase too when phase_names is given)."""
    from scipy.io import netcdf_file
    dt = np.float64 if double_precision else np.float32
    code = "d" if double_precision else "f"

    def as3(a):
        a = np.asarray(a, dtype=np.float64)
        return a[None] if a.ndim == 2 else a

    phases = [phase] if phase_names is None else list(phase)
    names = ["phase"] if phase_names is None else list(phase_names)
    nz, ny, nx = as3(phases[0]).shape
    f = netcdf_file(path, "w", version=2)
    try:
        f.createDimension("x", nx)
        f.createDimension("y", ny)
        f.createDimension("z", nz)
        f.createDimension("ns", 0 if conc is None else len(conc))
        if quat is not None:
            f.createDimension("qlen", len(quat))
        for s, cs in enumerate([] if conc is None else conc):
            v = f.createVariable(f"concentration{s}", code, ("z", "y", "x"))
            v[:] = as3(cs).astype(dt)
        if temperature is not None:
            v = f.createVariable("temperature", code, ("z", "y", "x"))
            v[:] = as3(temperature).astype(dt)
        for nm, a in zip(names, phases):
            v = f.createVariable(nm, code, ("z", "y", "x"))
            v[:] = as3(a).astype(dt)
        if eta is not None:
            v = f.createVariable("eta", code, ("z", "y", "x"))
            v[:] = as3(eta).astype(dt)
        for m, qm in enumerate([] if quat is None else quat):
            v = f.createVariable(f"quat{m + 1}", code, ("z", "y", "x"))
            v[:] = as3(qm).astype(dt)
    finally:
        f.close()


def patch_values(a, boxes, ndim):
    """Interior values of a domain array (with or without a leading component
    axis) in a level's compact order, patch by patch: what each patch's
    initializePatchFromData reads from its slab."""
    parts = []
    lead = a.ndim > ndim
    for b in boxes:
        if ndim == 3:
            sl = (slice(b.lo[2], b.hi[2] + 1), slice(b.lo[1], b.hi[1] + 1),
                  slice(b.lo[0], b.hi[0] + 1))
        else:
            sl = (slice(b.lo[1], b.hi[1] + 1), slice(b.lo[0], b.hi[0] + 1))
        blk = a[(slice(None),) + sl] if lead else a[sl]
        parts.append(blk.reshape(a.shape[0], -1) if lead else blk.reshape(-1))
    return np.concatenate(parts, axis=-1)
