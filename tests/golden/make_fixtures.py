"""Extract small golden fixtures from the reference's own test data.

Run in the development container only (it reads the reference checkout):

    python tests/golden/make_fixtures.py /root/reference/testdata

It decodes the serialized kernel-argument dumps that the reference's replay
test uses (opencl/tests/main.cpp:110-340, format written by
btLauncherCL::serializeToFile, opencl/broadphase_benchmark/btLauncherCL.h:253-278)
and writes compact .npz fixtures next to this script:

  sap_aabbs.npz    5,876 world AABBs from m_sapKernelArgs.bin (a real demo
                   scene); input of the broadphase known-answer test.
  clip_golden.npz  380 convex pairs from clipFacesAndContactReductionKernel
                   _before/_after.bin: world-space reference faces, incident
                   polygons, separating normals, and the contacts the
                   reference kernel produced (the only golden input/output pair
                   the reference holds).

m_batchSolveKernel.7z is not extracted: it is input-only (no gold output) and
its 96-byte body records belong to an older layout than the pipeline's.

Only data is written; no reference source is copied.  The archives are 7z
(LZMA) and zip; a minimal 7z container reader is implemented here because
py7zr is not installed.
"""
import lzma
import os
import struct
import sys
import zipfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------- 7z reader
def _num(b, i):
    first = b[i]
    i += 1
    mask = 0x80
    val = 0
    for n in range(8):
        if first & mask == 0:
            val |= (first & (mask - 1)) << (8 * n)
            return val, i
        val |= b[i] << (8 * n)
        i += 1
        mask >>= 1
    return val, i


def _streams(b, i):
    info = {}
    while True:
        t = b[i]
        i += 1
        if t == 0x00:
            return info, i
        if t == 0x06:
            info["packpos"], i = _num(b, i)
            n, i = _num(b, i)
            sizes = []
            while True:
                t2 = b[i]
                i += 1
                if t2 == 0x00:
                    break
                if t2 == 0x09:
                    for _ in range(n):
                        s, i = _num(b, i)
                        sizes.append(s)
                elif t2 == 0x0A:
                    assert b[i] == 1
                    i += 1 + 4 * n
            info["packsizes"] = sizes
        elif t == 0x07:
            assert b[i] == 0x0B
            i += 1
            nf, i = _num(b, i)
            assert b[i] == 0
            i += 1
            folders = []
            for _ in range(nf):
                nc, i = _num(b, i)
                coders = []
                for _ in range(nc):
                    flag = b[i]
                    i += 1
                    cid = b[i:i + (flag & 0xF)]
                    i += flag & 0xF
                    props = b""
                    if flag & 0x20:
                        ps, i = _num(b, i)
                        props = b[i:i + ps]
                        i += ps
                    coders.append((cid, props))
                folders.append(coders)
            assert b[i] == 0x0C
            i += 1
            unpack = []
            for _ in folders:
                s, i = _num(b, i)
                unpack.append(s)
            if b[i] == 0x0A:
                i += 1
                assert b[i] == 1
                i += 1 + 4 * nf
            assert b[i] == 0x00
            i += 1
            info["folders"], info["unpack"] = folders, unpack
        elif t == 0x08:
            while True:
                t2 = b[i]
                i += 1
                if t2 == 0x00:
                    break
                if t2 == 0x0D:
                    info["numsub"] = []
                    for _ in info["folders"]:
                        n, i = _num(b, i)
                        info["numsub"].append(n)
                elif t2 == 0x09:
                    sizes = []
                    for fi, n in enumerate(info.get("numsub", [1] * len(info["folders"]))):
                        tot = 0
                        for _ in range(n - 1):
                            s, i = _num(b, i)
                            sizes.append(s)
                            tot += s
                        sizes.append(info["unpack"][fi] - tot)
                    info["subsizes"] = sizes
                elif t2 == 0x0A:
                    tot = sum(info.get("numsub", [1]))
                    assert b[i] == 1
                    i += 1 + 4 * tot
        else:
            raise ValueError("unexpected 7z property 0x%02x" % t)


def _lzma(data, props, size):
    p0 = props[0]
    lc, p0 = p0 % 9, p0 // 9
    lp, pb = p0 % 5, p0 // 5
    ds = struct.unpack("<I", props[1:5])[0]
    dec = lzma.LZMADecompressor(format=lzma.FORMAT_RAW, filters=[
        {"id": lzma.FILTER_LZMA1, "dict_size": ds, "lc": lc, "lp": lp, "pb": pb}])
    return dec.decompress(data, max_length=size)[:size]


def _folder(d, info, fi):
    off = 32 + info["packpos"] + sum(info["packsizes"][:fi])
    (cid, props), = info["folders"][fi]
    assert cid == b"\x03\x01\x01", "only LZMA folders are supported"
    return _lzma(d[off:off + info["packsizes"][fi]], props, info["unpack"][fi])


def read_7z(path):
    d = open(path, "rb").read()
    assert d[:6] == b"7z\xbc\xaf\x27\x1c"
    nho, nhs, _ = struct.unpack("<QQI", d[12:32])
    hdr = d[32 + nho:32 + nho + nhs]
    if hdr[0] == 0x17:
        info, _ = _streams(hdr, 1)
        hdr = _folder(d, info, 0)
    assert hdr[0] == 0x01 and hdr[1] == 0x04
    info, i = _streams(hdr, 2)
    assert hdr[i] == 0x05
    i += 1
    _, i = _num(hdr, i)
    names = []
    while True:
        t = hdr[i]
        i += 1
        if t == 0:
            break
        sz, i = _num(hdr, i)
        if t == 0x11:
            names = [s for s in hdr[i + 1:i + sz].decode("utf-16-le").split("\x00") if s]
        i += sz
    blob = b"".join(_folder(d, info, fi) for fi in range(len(info["folders"])))
    out, o = {}, 0
    for n, s in zip(names, info.get("subsizes", info["unpack"])):
        out[n] = blob[o:o + s]
        o += s
    return out


# ------------------------------------------------------ kernel-arg records
def parse_args(buf):
    """btKernelArgData records as written by a 32-bit build: int isBuffer,
    int argIndex, int sizeInBytes, 16-byte union (28 bytes), each buffer
    followed by its bytes; the work-item count trails the record list."""
    n = struct.unpack_from("<i", buf, 0)[0]
    i, args = 4, []
    for _ in range(n):
        isb, idx, sz = struct.unpack_from("<iii", buf, i)
        i += 12
        raw = buf[i:i + 16]
        i += 16
        if isb:
            args.append(buf[i:i + sz])
            i += sz
        else:
            args.append(raw[:sz])
    trailer = struct.unpack_from("<i", buf, i)[0]
    return args, trailer


def f32(b, *shape):
    return np.frombuffer(b, dtype=np.float32).reshape(*shape).copy()


def i32(b, *shape):
    return np.frombuffer(b, dtype=np.int32).reshape(*shape).copy()


def make_sap(src):
    buf = zipfile.ZipFile(os.path.join(src, "m_sapKernelArgs.zip")).read("m_sapKernelArgs.bin")
    args, nwork = parse_args(buf)
    n = struct.unpack("<i", args[3])[0]
    assert n == nwork == 5876
    aabb = f32(args[0], n, 8)
    handles = i32(args[0], n, 8)[:, 3]
    np.savez_compressed(os.path.join(HERE, "sap_aabbs.npz"), aabb_min=aabb[:, 0:3],
                        aabb_max=aabb[:, 4:7], handle=handles)
    print("sap_aabbs.npz: %d AABBs" % n)


def make_clip(src):
    files = read_7z(os.path.join(src, "clipFacesAndContactReductionKernel.7z"))
    before, nb = parse_args(files["clipFacesAndContactReductionKernel_before.bin"])
    after, na = parse_args(files["clipFacesAndContactReductionKernel_after.bin"])
    cap = struct.unpack("<i", before[11])[0]
    npairs = struct.unpack("<i", before[12])[0]
    assert nb == na == npairs == 380 and cap == 64
    pairs = i32(before[0], -1, 2)[:npairs]
    sep = f32(before[2], -1, 4)[:npairs]
    has_sep = i32(before[3], -1)[:npairs]
    faces = i32(before[5], -1, 4)[:npairs]
    vA = f32(before[6], -1, cap, 4)[:npairs]
    nA = f32(before[7], -1, 4)[:npairs]
    vB = f32(before[8], -1, cap, 4)[:npairs]
    ncont = struct.unpack("<i", after[10])[0]
    craw = np.frombuffer(after[4], dtype=np.uint8).reshape(-1, 96)[:ncont]
    cpos = craw[:, :64].copy().view(np.float32).reshape(-1, 4, 4)
    cnrm = craw[:, 64:80].copy().view(np.float32).reshape(-1, 4)
    cint = craw[:, 80:96].copy().view(np.int32).reshape(-1, 4)
    kA = int(faces[:, 2].max())
    kB = int(faces[:, 3].max())
    np.savez_compressed(
        os.path.join(HERE, "clip_golden.npz"), pairs=pairs, sep=sep, has_sep=has_sep,
        faces=faces, face_a=vA[:, :kA], normal_a=nA, poly_b=vB[:, :kB],
        contact_pos=cpos, contact_normal=cnrm, contact_int=cint,
        after_b1=f32(after[8], -1, cap, 4)[:npairs, :16],
        after_b2=f32(after[9], -1, cap, 4)[:npairs, :16])
    print("clip_golden.npz: %d pairs, %d contacts" % (npairs, ncont))


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/testdata"
    make_sap(src)
    make_clip(src)
