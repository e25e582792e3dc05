"""Sigma (spectral) files of the GFS: read, write and conversion to the model
state, for initial conditions and restarts.

File format (src/sigio_module.f90): Fortran sequential unformatted records,
big-endian (sigio_sropen/swopen open with convert='big_endian', :437,451),
each record framed by its 4-byte length.

* Modern header, ivs = 200509 (sigio_srhead :482-512, sigio_swhead :589-613):
  'GFS SIG ' + ivs, nhead, ndata, 3 reserved; lhead(nhead), ldata(ndata);
  a 50-word record (fhour, idate(4), jcap, levs, itrun, iorder, irealf, igen,
  latf, lonf, latb, lonb, latr, lonr, ntrac, nvcoord, icen2, iens(2), idpp,
  idsl, idvc, idvm, idvt, idrun, idusr, pdryini, ncldt, ixgr, 18 zeros);
  vcoord(levs+1, nvcoord); cfvars(5+ntrac) as 8-character names; with
  idvm / 10 % 10 == 3 also cpi(ntrac+1) and ri(ntrac+1). Record lengths as
  sigio_adhead (:1345-1371). The reader, like sigio_srhead, returns pdryini
  as the fill value -9999 for this version.
* Legacy header, ivs = 198410 (:518-567, :614-672): a 32-character label, then
  fhour, idate(4), sisl(201), ext(44) (reals), with si / sl (idvc 0, 1) or ak /
  bk (idvc 2, 3) in sisl and the integer header fields in ext.
* Data (sigio_srdata :799-833, sigio_swdata): hs, ps, t(levs), (d, z)(levs),
  q(levs, ntrac), then xgr / xss when ixgr asks for them; each record
  (jcap+1)(jcap+2) reals (4-byte, or 8-byte when irealf = 2) holding the
  complex coefficients as (re, im) pairs in the model's m-major order.

Conversion to the model (src/dyn_init.f90): coefficients times sqrt(2 pi)
(copyspecin :93-104, NCEP to orthonormal normalisation) or divided by it on
output (copyspecout :106-117); ln(ps) in cb in the file, in Pa in the model,
converted in grid space (:48-51 on input, wrtout_sig :156-159 on output);
ak / bk the header's vcoord reversed to top -> bottom (:52-57); the restart
time fhour * 3600 s and pdryini * 1000 Pa, reset to "compute on the first
step" when it is not in (0, 1e5] Pa (params.f90:255-264; the reference keeps
a non-positive value, which its mass fixer cannot use).
"""
import struct

import numpy as np

FILL = -9999.0
SQRT2PI = np.sqrt(2.0 * np.pi)
LHEAD1 = 32
_INTS = ("jcap", "levs", "itrun", "iorder", "irealf", "igen", "latf", "lonf", "latb", "lonb", "latr", "lonr",
         "ntrac", "nvcoord", "icen2")
_INTS2 = ("idpp", "idsl", "idvc", "idvm", "idvt", "idrun", "idusr")
# legacy ext(1:26) order (sigio_srhead :521-546)
_EXT = ("jcap", "levs", "itrun", "iorder", "irealf", "igen", "lonf", "latf", "lonb", "latb", "lonr", "latr",
        "ntrac", "icen2", "iens1", "iens2", "idpp", "idsl", "idvc", "idvm", "idvt", "idrun", "idusr", "pdryini",
        "ncldt", "ixgr")


class SigioError(RuntimeError):
    pass


# ---- Fortran sequential records (big-endian) --------------------------------------

def _records(buf):
    out, o = [], 0
    while o < len(buf):
        if o + 4 > len(buf):
            raise SigioError("truncated record marker")
        n = struct.unpack(">i", buf[o:o + 4])[0]
        if n < 0 or o + 8 + n > len(buf) or struct.unpack(">i", buf[o + 4 + n:o + 8 + n])[0] != n:
            raise SigioError(f"bad record framing at byte {o}")
        out.append(buf[o + 4:o + 4 + n])
        o += 8 + n
    return out


def _rec(payload):
    return struct.pack(">i", len(payload)) + payload + struct.pack(">i", len(payload))


def _nxgr(head):
    """grid fields and surface scalars that ixgr adds (sigio_adhead :1327-1344)"""
    return {1: (2 * head["levs"] + 1, 0), 2: (4 * head["levs"] + 3, 0), 3: (3 * head["levs"] + 1, 0),
            4: (4 * head["levs"] + 3, 1), 5: (3 * head["levs"] + 1, 1)}.get(head.get("ixgr", 0), (0, 0))


def header_lengths(head):
    """lhead and ldata of sigio_adhead (:1345-1371)"""
    levs, ntrac, nvc = head["levs"], head["ntrac"], head.get("nvcoord", 2)
    nxgr, nxss = _nxgr(head)
    nspec = 2 + (3 + ntrac) * levs
    irealf = head.get("irealf", 1)
    nc = (head["jcap"] + 1) * (head["jcap"] + 2)
    if head["ivs"] == 200509:
        thermo3 = (head.get("idvm", 0) // 10) % 10 == 3
        nhead = 7 if thermo3 else 5
        ndata = nspec + nxgr + (1 if nxss > 0 else 0)
        lhead = [LHEAD1, 4 * (nhead + ndata), 4 * 50, 4 * ((levs + 1) * nvc), 8 * (5 + ntrac)]
        if thermo3:
            lhead += [4 * (ntrac + 1), 4 * (ntrac + 1)]
        ldata = [4 * irealf * nc] * nspec + [4 * irealf * head["lonb"] * head["latb"]] * nxgr
        ldata += [4 * irealf * nxss] * (ndata - nspec - nxgr)
        return lhead, ldata
    ldata = [4 * irealf * nc] * nspec + [4 * irealf * head["lonb"] * head["latb"]] * nxgr
    return [LHEAD1, 4 * 250], ldata


# ---- read -----------------------------------------------------------------------

def read(path):
    """(header dict, data dict) of a sigma file. Data arrays (float32, or
    float64 when irealf = 2): hs, ps (nc), t, d, z (levs, nc), q (ntrac, levs,
    nc), xgr (nxgr, latb, lonb), xss (nxss); nc = (jcap+1)(jcap+2) reals."""
    with open(path, "rb") as f:
        recs = _records(f.read())
    if not recs:
        raise SigioError("empty file")
    head = {}
    if recs[0][:8] == b"GFS SIG ":
        _, _, ivs, nhead, ndata, *_ = struct.unpack(">4s4s6i", recs[0][:32])
        if ivs != 200509:
            raise SigioError(f"unsupported sigma file version {ivs}")
        head.update(ivs=ivs, nhead=nhead, ndata=ndata, clabsig=" ")
        w = recs[2]
        head["fhour"] = struct.unpack(">f", w[0:4])[0]
        head["idate"] = list(struct.unpack(">4i", w[4:20]))
        v = struct.unpack(">15i", w[20:80])
        head.update(dict(zip(_INTS, v)))
        head["iens"] = list(struct.unpack(">2i", w[80:88]))
        head.update(dict(zip(_INTS2, struct.unpack(">7i", w[88:116]))))
        head["pdryini_record"] = struct.unpack(">f", w[116:120])[0]
        head["ncldt"], head["ixgr"] = struct.unpack(">2i", w[120:128])
        levs, nvc = head["levs"], head["nvcoord"]
        head["vcoord"] = np.frombuffer(recs[3], ">f4").astype(np.float32).reshape(nvc, levs + 1).T.copy()
        head["cfvars"] = [recs[4][8 * i:8 * i + 8].decode() for i in range(len(recs[4]) // 8)]
        nh = 5
        if (head["idvm"] // 10) % 10 == 3:
            head["cpi"] = np.frombuffer(recs[5], ">f4").astype(np.float32)
            head["ri"] = np.frombuffer(recs[6], ">f4").astype(np.float32)
            nh = 7
        # sigio_srhead :507-512: the label, si, sl, ak, bk and pdryini are reset for this version
        head["pdryini"] = FILL
        data_recs = recs[nh:]
    else:
        head["clabsig"] = recs[0][:LHEAD1].decode(errors="replace")
        w = recs[1]
        head["fhour"] = struct.unpack(">f", w[0:4])[0]
        head["idate"] = list(struct.unpack(">4i", w[4:20]))
        sisl = np.frombuffer(w[20:20 + 4 * 201], ">f4").astype(np.float32)
        ext = np.frombuffer(w[20 + 4 * 201:20 + 4 * 245], ">f4").astype(np.float64)
        e = dict(zip(_EXT, ext))
        for k in _EXT:
            if k not in ("pdryini", "iens1", "iens2"):
                head[k] = int(e[k])
        head["iens"] = [int(e["iens1"]), int(e["iens2"])]
        head["ntrac"] = int(max(e["ntrac"], 1.0))
        head["pdryini"] = float(np.float32(e["pdryini"]))
        levs = head["levs"]
        if head["idvc"] in (0, 1):
            head["nvcoord"] = 1
            head["si"] = sisl[:levs + 1].copy()
            head["sl"] = sisl[levs + 1:2 * levs + 1].copy()
            head["vcoord"] = sisl[:levs + 1].reshape(levs + 1, 1).copy()
        elif head["idvc"] in (2, 3):
            head["nvcoord"] = 2
            head["vcoord"] = np.stack([sisl[:levs + 1], sisl[levs + 1:2 * levs + 2]], 1).copy()
        else:
            raise SigioError(f"unknown vertical coordinate type {head['idvc']}")
        head["ivs"] = 198410
        data_recs = recs[2:]
    head["nxgr"], head["nxss"] = _nxgr(head)
    levs, ntrac = head["levs"], head["ntrac"]
    nc = (head["jcap"] + 1) * (head["jcap"] + 2)
    dt = ">f8" if head.get("irealf", 1) == 2 else ">f4"
    nspec = 2 + (3 + ntrac) * levs
    if len(data_recs) < nspec:
        raise SigioError(f"{len(data_recs)} data records, expected at least {nspec}")

    def arr(r, n):
        a = np.frombuffer(r, dt)
        if a.size < n:
            raise SigioError("short data record")
        return a[:n].astype(np.float64 if dt == ">f8" else np.float32)

    it = iter(data_recs)
    data = {"hs": arr(next(it), nc), "ps": arr(next(it), nc)}
    data["t"] = np.stack([arr(next(it), nc) for _ in range(levs)])
    d, z = [], []
    for _ in range(levs):
        d.append(arr(next(it), nc))
        z.append(arr(next(it), nc))
    data["d"], data["z"] = np.stack(d), np.stack(z)
    q = [np.stack([arr(next(it), nc) for _ in range(levs)]) for _ in range(ntrac)]
    data["q"] = np.stack(q) if q else np.zeros((0, levs, nc), np.float64 if dt == ">f8" else np.float32)
    ng = head.get("lonb", 0) * head.get("latb", 0)
    data["xgr"] = np.stack([arr(next(it), ng).reshape(head["latb"], head["lonb"])
                            for _ in range(head["nxgr"])]) if head["nxgr"] else np.zeros((0, 0, 0), np.float32)
    data["xss"] = arr(next(it), head["nxss"]) if head["nxss"] else np.zeros(0, np.float32)
    return head, data


# ---- write ----------------------------------------------------------------------

def write(path, head, data):
    """Write a sigma file in the header version head['ivs'] (200509 or 198410)."""
    head = dict(head)
    levs, ntrac = head["levs"], head["ntrac"]
    nc = (head["jcap"] + 1) * (head["jcap"] + 2)
    irealf = head.get("irealf", 1)
    dt = ">f8" if irealf == 2 else ">f4"
    nxgr, nxss = _nxgr(head)
    vc = np.asarray(head["vcoord"], np.float32).reshape(levs + 1, -1)
    head.setdefault("nvcoord", vc.shape[1])
    out = []
    if head["ivs"] >= 200509:
        head["ivs"] = 200509
        lhead, ldata = header_lengths(head)
        out.append(struct.pack(">4s4s6i", b"GFS ", b"SIG ", 200509, len(lhead), len(ldata), 0, 0, 0))
        out.append(struct.pack(f">{len(lhead) + len(ldata)}i", *lhead, *ldata))
        w = struct.pack(">f4i", head.get("fhour", 0.0), *head.get("idate", [0, 0, 0, 0]))
        w += struct.pack(">15i", *[int(head.get(k, 0)) for k in _INTS[:13]], head["nvcoord"], head.get("icen2", 0))
        w += struct.pack(">2i", *head.get("iens", [0, 0]))
        w += struct.pack(">7i", *[int(head.get(k, 0)) for k in _INTS2])
        w += struct.pack(">f2i", head.get("pdryini", 0.0), head.get("ncldt", 0), head.get("ixgr", 0))
        w += struct.pack(">18i", *([0] * 18))
        out.append(w)
        out.append(vc.T.astype(">f4").tobytes())
        cf = head.get("cfvars") or ["hs", "ps", "t", "d", "z"] + [f"q{i + 1}" for i in range(ntrac)]
        if len(cf) != 5 + ntrac:
            raise SigioError("cfvars must hold 5 + ntrac names")
        out.append(b"".join(c.encode()[:8].ljust(8) for c in cf))
        if (head.get("idvm", 0) // 10) % 10 == 3:
            out.append(np.asarray(head["cpi"], ">f4").tobytes())
            out.append(np.asarray(head["ri"], ">f4").tobytes())
    else:
        out.append(str(head.get("clabsig", "")).encode()[:LHEAD1].ljust(LHEAD1))
        sisl = np.zeros(201, np.float32)
        if head["idvc"] in (0, 1):
            sisl[:levs + 1] = head.get("si", vc[:, 0])
            if "sl" in head:
                sisl[levs + 1:2 * levs + 1] = head["sl"]
        else:
            sisl[:levs + 1] = vc[:, 0]
            sisl[levs + 1:2 * levs + 2] = vc[:, 1]
        e = {k: float(head.get(k, 0)) for k in _EXT}
        e["iens1"], e["iens2"] = (float(x) for x in head.get("iens", [0, 0]))
        e["pdryini"] = float(head.get("pdryini", 0.0))
        ext = np.zeros(44, np.float32)
        ext[:26] = [e[k] for k in _EXT]
        out.append(struct.pack(">f4i", head.get("fhour", 0.0), *head.get("idate", [0, 0, 0, 0])) +
                   sisl.astype(">f4").tobytes() + ext.astype(">f4").tobytes())

    def rec(a, n):
        a = np.asarray(a).ravel()
        if a.size < n:
            raise SigioError("short data array")
        return a[:n].astype(dt).tobytes()

    out.append(rec(data["hs"], nc))
    out.append(rec(data["ps"], nc))
    out += [rec(data["t"][k], nc) for k in range(levs)]
    for k in range(levs):
        out += [rec(data["d"][k], nc), rec(data["z"][k], nc)]
    out += [rec(data["q"][n][k], nc) for n in range(ntrac) for k in range(levs)]
    ng = head.get("lonb", 0) * head.get("latb", 0)
    out += [rec(data["xgr"][n], ng) for n in range(nxgr)]
    if nxss:
        out.append(rec(data["xss"], nxss))
    with open(path, "wb") as f:
        f.write(b"".join(_rec(p) for p in out))


# ---- model state <-> file -----------------------------------------------------------

def spec_in(r):
    """copyspecin (dyn_init.f90:93-104): (re, im) pairs in NCEP normalisation ->
    complex orthonormal coefficients, on the last axis"""
    r = np.asarray(r, np.float64)
    return SQRT2PI * (r[..., 0::2] + 1j * r[..., 1::2])


def spec_out(c):
    """copyspecout (dyn_init.f90:106-117): complex orthonormal -> (re, im)
    pairs in NCEP normalisation, float32"""
    c = np.asarray(c)
    out = np.empty(c.shape[:-1] + (2 * c.shape[-1],), np.float32)
    out[..., 0::2] = (c.real / SQRT2PI).astype(np.float32)
    out[..., 1::2] = (c.imag / SQRT2PI).astype(np.float32)
    return out


def model_state(head, data, sht):
    """The model's spectral state from a sigma file, as readin_sig and init_dyn
    (dyn_init.f90:43-57) and params.f90:255-264 form it. sht supplies
    spectogrd / grdtospec on (field, ndimspec) complex and (field, nlat, nlon)
    real arrays at the file's grid. Returns vrt, div, tv (levs, ndimspec),
    trac (ntrac, levs, ndimspec), lnps (ln Pa), topo, ak, bk (top -> bottom,
    Pa / unitless), tstart (s) and pdryini (Pa, 0: compute on the first step)."""
    if head["idvc"] != 2:
        raise SigioError(f"unknown vertical coordinate type {head['idvc']}")
    levs = head["levs"]
    lnps_cb = spec_in(data["ps"])
    psg = 1000.0 * np.exp(sht.spectogrd(lnps_cb[None])[0])      # cb -> Pa
    lnps = sht.grdtospec(np.log(psg)[None])[0]
    vc = np.asarray(head["vcoord"], np.float64)
    pd = 1000.0 * float(head.get("pdryini", 0.0))
    if not (0.0 < pd <= 1.0e5):
        pd = 0.0
    return dict(vrt=spec_in(data["z"]), div=spec_in(data["d"]), tv=spec_in(data["t"]),
                trac=spec_in(data["q"]), lnps=lnps, topo=spec_in(data["hs"]),
                ak=vc[::-1, 0].copy(), bk=vc[::-1, 1].copy(), tstart=3600.0 * float(head["fhour"]),
                pdryini=pd)


def sigma_data(state, sht, ntrac=None):
    """wrtout_sig (dyn_init.f90:145-177): the file's data records from the
    model state (dict as model_state returns; ln ps in Pa -> ln ps in cb in
    grid space)"""
    psg = np.exp(sht.spectogrd(np.asarray(state["lnps"])[None])[0]) / 1000.0  # Pa -> cb
    lnps_cb = sht.grdtospec(np.log(psg)[None])[0]
    trac = np.asarray(state["trac"])
    if ntrac is not None:
        trac = trac[:ntrac]
    return dict(hs=spec_out(state["topo"]), ps=spec_out(lnps_cb), t=spec_out(state["tv"]),
                d=spec_out(state["div"]), z=spec_out(state["vrt"]), q=spec_out(trac),
                xgr=np.zeros((0, 0, 0), np.float32), xss=np.zeros(0, np.float32))


def header_for(nlon, nlat, ntrunc, ak, bk, ntrac, fhour=0.0, pdryini=0.0, idate=(0, 1, 1, 2000)):
    """A modern (ivs 200509) hybrid-coordinate header for the model's grid;
    ak, bk top -> bottom (stored bottom -> top in vcoord, as dyn_init.f90:53-56
    reads them), pdryini in Pa (stored in cb, wrtout_sig :168)."""
    ak, bk = np.asarray(ak, np.float64), np.asarray(bk, np.float64)
    levs = ak.size - 1
    return dict(ivs=200509, fhour=float(fhour), idate=list(idate), jcap=ntrunc, levs=levs, itrun=1, iorder=2,
                irealf=1, igen=0, latf=nlat, lonf=nlon, latb=nlat, lonb=nlon, latr=nlat, lonr=nlon, ntrac=ntrac,
                nvcoord=2, icen2=0, iens=[0, 0], idpp=0, idsl=0, idvc=2, idvm=0, idvt=0, idrun=0, idusr=0,
                pdryini=pdryini / 1000.0, ncldt=0, ixgr=0,
                vcoord=np.stack([ak[::-1], bk[::-1]], 1).astype(np.float32))
