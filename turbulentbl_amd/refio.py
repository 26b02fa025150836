"""The reference's on-disk formats, read and written byte-compatibly.

* Flow field dumps.
  * Combined restart file `Input/3D.dat`, read at `dnssetbl/dnssetbl.f90:118-135` and written by
    `dnssetbl/combineFile.f90:55-71`. It holds every span row j = 1..nY.
  * Per-rank dump `Output/3DNNN.dat` (`dnssetbl.f90:907-934`). It holds rows j = 2..nY-1, and the
    header is on rank 1's file only.
* Recycling statistics `Input/statData.dat` (`dnssetbl.f90:137-142,1193-1197`).
* Inflow mean profile `Input/initProf.dat` (`dnssetbl.f90:106-111`).
* The statistics files written at every time-average rescale
  (`dnssetbl.f90:1071-1198`): `Output/lamda.dat` (appended), `Output/W.dat`,
  `Output/boundLyrThik.dat`, `Output/newK.dat`, `Output/RMS.dat` and
  `Output/meanVel.dat`, plus `Input/statData.dat`.

Records use the reference's FORMAT statements: 101/105/107/108 for the headers,
104 for the data, 115 and 103. Numbers are written the way Fortran's E, F and I
edit descriptors write them (`Ew.d` mantissa 0.d1d2...E+xx).
"""
import math

import numpy as np


def _E(x, w=22, d=15):
    """Fortran Ew.d (d significant digits, mantissa in [0.1, 1))"""
    x = float(x)
    if x == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m = "%.*e" % (d - 1, abs(x))                # d significant digits, correctly rounded
        digits, e = m.split("e")
        e10 = int(e) + 1
        s = "0." + digits.replace(".", "") + ("E%+03d" % e10 if abs(e10) < 100 else "%+04d" % e10)
        if x < 0:
            s = "-" + s
    return s.rjust(w)


def _F(x, w, d):
    """Fortran Fw.d: NaN and Inf spelled as Fortran writes them, a field of
    asterisks when the value does not fit"""
    x = float(x)
    if x != x:
        return "NaN".rjust(w)
    if x in (math.inf, -math.inf):
        return ("Inf" if x > 0 else "-Inf").rjust(w)
    s = "%*.*f" % (w, d, x)
    return "*" * w if len(s) > w else s


def _I(n, w, m=1):
    return ("%0*d" % (m, int(n))).rjust(w)


def record104(x1, x2, x3, u, v, w, p):
    """FORMAT(3(1X,F10.6),4(1X,E22.15))"""
    return "".join(" " + _F(a, 10, 6) for a in (x1, x2, x3)) + "".join(" " + _E(a) for a in (u, v, w, p))


def write_3d(path, x1, x2, x3, Ut, Press, time, nstep):
    """combined restart file (all j = 1..nY), as combineFile writes it"""
    nX, nY, nZ = Press.shape
    with open(path, "w") as f:
        f.write('TITLE ="' + _F(time, 20, 10) + _I(nstep, 9) + '"\n')                      # FORMAT 105
        f.write('Variables = "x","y","z","U","V","W","P"\n')
        f.write("ZONE k=" + _I(nX, 4) + ",j=" + _I(nY, 4) + ",i=" + _I(nZ, 4) + ",DATAPACKING=POINT\n")
        for i in range(nX):
            f.write("\n")
            for j in range(nY):
                f.write("\n")
                for k in range(nZ):
                    f.write(record104(x1[i], x2[j], x3[k], Ut[i, j, k, 0], Ut[i, j, k, 1], Ut[i, j, k, 2],
                                      Press[i, j, k]) + "\n")


def write_3d_rank(path, x1, x2, x3, Ut, Press, time, nstep, nproc=1, kb=1, ke=None, header=True):
    """per-rank dump Output/3DNNN.dat: rows j = 2..nY-1, planes kb..ke (1-based)"""
    nX, nY, nZ = Press.shape
    ke = nZ if ke is None else ke
    with open(path, "w") as f:
        if header:
            f.write('TITLE ="' + _F(time, 20, 10) + _I(nstep, 9) + _I(nproc, 3) + '"\n')     # FORMAT 107
            f.write('Variables = "x","y","z","U","V","W","P"\n')
            f.write("ZONE k=" + _I(nX, 4) + ",j=" + _I(nY - 2, 4) + ",i=" + _I(nZ, 4) + ',DATAPACKING="POINT"\n')
        for i in range(nX):
            f.write("\n")
            for j in range(1, nY - 1):
                f.write("\n")
                for k in range(kb - 1, ke):
                    f.write(record104(x1[i], x2[j], x3[k], Ut[i, j, k, 0], Ut[i, j, k, 1], Ut[i, j, k, 2],
                                      Press[i, j, k]) + "\n")


def read_3d(path, nX, nY, nZ):
    """restart read of dnssetbl.f90:118-133 -> time, nstep, x1, x2, x3, Ut, Press"""
    Ut = np.zeros((nX, nY, nZ, 3), order="F")
    Press = np.zeros((nX, nY, nZ), order="F")
    x1, x2, x3 = np.zeros(nX), np.zeros(nY), np.zeros(nZ)
    with open(path) as f:
        head = f.readline()
        time = float(head[8:28])                    # FORMAT(8X,F20.10,I9,1X)
        nstep = int(head[28:37])
        f.readline()
        f.readline()
        for i in range(nX):
            f.readline()
            for j in range(nY):
                f.readline()
                for k in range(nZ):
                    v = [float(t) for t in f.readline().split()]
                    x1[i], x2[j], x3[k] = v[0], v[1], v[2]
                    Ut[i, j, k, :] = v[3:6]
                    Press[i, j, k] = v[6]
    return time, nstep, x1, x2, x3, Ut, Press


def write_stat(path, u, v, w, wt, newk):
    """Input/statData.dat: usum2(68,k) vsum2(68,k) wsum2(68,k) Wt(1,k) new(k), FORMAT(4(1X,F15.7),I5.3)"""
    with open(path, "w") as f:
        for k in range(len(u)):
            f.write("".join(" " + _F(a, 15, 7) for a in (u[k], v[k], w[k], wt[k])) + _I(newk[k], 5, 3) + "\n")


def read_stat(path, nZ):
    u, v, w, wt = (np.zeros(nZ) for _ in range(4))
    newk = np.zeros(nZ, dtype=np.int32)
    with open(path) as f:
        for k in range(nZ):
            line = f.readline()
            vals = [float(line[1 + 16 * q:16 + 16 * q]) for q in range(4)]
            u[k], v[k], w[k], wt[k] = vals
            newk[k] = int(line[64:69])
    return u, v, w, wt, newk


def write_init_prof(path, ume, lead=None):
    """Input/initProf.dat: 24 leading columns (free text) then E18.8 (FORMAT(24x,E18.8) on read)"""
    with open(path, "w") as f:
        for k, val in enumerate(ume):
            tag = ("" if lead is None else lead[k]).ljust(24)[:24]
            f.write(tag + _E(val, 18, 8) + "\n")


def read_init_prof(path, n=37):
    out = np.zeros(n)
    with open(path) as f:
        for k in range(n):
            out[k] = float(f.readline()[24:42])
    return out


def _list_header(text):
    """WRITE(u,*) 'text': list-directed output of one character constant
    (one leading blank)"""
    return " " + text + "\n"


def rescale_eta(Wt_rec, lamda, lamda03):
    """eta(k), k = 1..nZ-1, of dnssetbl.f90:1096-1105 (lamda where the Lund
    weight at the recycling plane is <= 0.99, else lamda03); eta(nZ) is never
    assigned by the reference and is written as 0"""
    nZ = len(Wt_rec)
    eta = np.zeros(nZ)
    for k in range(nZ - 1):
        eta[k] = lamda if Wt_rec[k] <= 0.99 else lamda03
    return eta


def reynolds_stresses(avg):
    """Ruu, Rvv, Rww, Ruv, Ruw, Rvw of dnssetbl.f90:1044-1053 from the nine
    averages usum2, vsum2, wsum2, uusum2, vvsum2, wwsum2, uvsum2, uwsum2,
    vwsum2 (avg[..., 0..8], each (nX, nZ))"""
    u, v, w, uu, vv, ww, uv, uw, vw = (avg[..., q] for q in range(9))
    return (uu - u * u, vv - v * v, ww - w * w, uv - u * v, uw - u * w, vw - v * w)


def write_rescale_outputs(outdir, x1, x3, avg, Wt, newk, bltkns, deltastar, theta, lamda, lamda03, time, Ue, Re,
                          i_rec, statdata=None):
    """The files of one statistics rescale, in the reference's formats:

    * `lamda.dat`, appended: FORMAT 110 (3(2x,F20.10)), time lamda lamda03;
    * `W.dat`: x1 x3 Wt, FORMAT 111 (3(1X,F15.7));
    * `boundLyrThik.dat`: x1 delta99 delta* theta H Re_theta Re_x, FORMAT 112
      (8(2x,F20.10)); Re_x is never assigned by the reference and is 0;
    * `newK.dat`: k new(k) eta(k), FORMAT 113 (2(I5),F20.10);
    * `RMS.dat`: x1 x3 Ruu Rvv Rww Ruv Ruw Rvw, FORMAT 116 (8(1X,F15.7));
    * `meanVel.dat`: x1 x3 usum2 vsum2 wsum2, FORMAT 114 (5(1X,F15.7));
    * `statdata` (the path of Input/statData.dat) if given: write_stat.

    avg is (nX, nZ, 9); i_rec is 1-based."""
    import os
    nX, nZ = len(x1), len(x3)
    r = i_rec - 1
    with open(os.path.join(outdir, "lamda.dat"), "a") as f:
        f.write("".join("  " + _F(a, 20, 10) for a in (time, lamda, lamda03)) + "\n")
    with open(os.path.join(outdir, "W.dat"), "w") as f:
        f.write(_list_header('Variables = "x1","x3","Wt"'))
        for i in range(nX):
            for k in range(nZ):
                f.write("".join(" " + _F(a, 15, 7) for a in (x1[i], x3[k], Wt[i, k])) + "\n")
    H = deltastar / theta
    Reth = theta * Re * Ue
    with open(os.path.join(outdir, "boundLyrThik.dat"), "w") as f:
        f.write(_list_header('Variables = "x1","BLthick","deltastar","theta","Ht","Reth","Rex"'))
        for i in range(nX):
            f.write("".join("  " + _F(a, 20, 10)
                            for a in (x1[i], bltkns[i], deltastar[i], theta[i], H[i], Reth[i], 0.0)) + "\n")
    eta = rescale_eta(Wt[r, :], lamda, lamda03)
    with open(os.path.join(outdir, "newK.dat"), "w") as f:
        f.write(_list_header('Variables = "k","newk","etak"'))
        for k in range(nZ):
            f.write(_I(k + 1, 5) + _I(newk[k], 5) + _F(eta[k], 20, 10) + "\n")
    R = reynolds_stresses(avg)
    with open(os.path.join(outdir, "RMS.dat"), "w") as f:
        f.write(_list_header('Variables = "x1","x3","Ruu","Rvv","Rww","Ruv","Ruw","Rvw"'))
        for i in range(nX):
            for k in range(nZ):
                f.write("".join(" " + _F(a, 15, 7) for a in (x1[i], x3[k]) + tuple(q[i, k] for q in R)) + "\n")
    with open(os.path.join(outdir, "meanVel.dat"), "w") as f:
        f.write(_list_header('Variables = "x1","x3","usum2","vsum2","wsum2"'))
        for i in range(nX):
            for k in range(nZ):
                f.write("".join(" " + _F(a, 15, 7) for a in (x1[i], x3[k], avg[i, k, 0], avg[i, k, 1], avg[i, k, 2]))
                        + "\n")
    if statdata is not None:
        write_stat(statdata, avg[r, :, 0], avg[r, :, 1], avg[r, :, 2], Wt[0, :], newk)


def read_table(path, header=True):
    """numeric columns of one of the files above (header line skipped)"""
    with open(path) as f:
        lines = f.readlines()[1 if header else 0:]
    return np.array([[float(t) for t in l.split()] for l in lines if l.strip()])
