"""Per-kernel resource figures (VGPRs, AGPRs, scratch, LDS) of the gfx950
code objects inside the built libqcu.so, read from their AMDGPU metadata.

Everything is parsed in this process, with no external tool: starting a
child process from a test process that has loaded the HIP runtime upsets
the BLAS thread pool that later tests use. The library's .hip_fatbin
section holds one offload bundle per translation unit (magic, entry count,
then offset / size / triple per entry), each with one gfx950 ELF whose
NT_AMDGPU_METADATA note (owner "AMDGPU", type 32) is a msgpack map with an
"amdhsa.kernels" list."""
import struct

import msgpack

MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
NT_AMDGPU_METADATA = 32
SHT_NOTE = 7
FIELDS = ("agpr_count", "vgpr_count", "sgpr_count", "private_segment_fixed_size", "group_segment_fixed_size")


def elf_sections(img):
    """[(name, type, bytes)] of a little-endian ELF64 image."""
    if img[:4] != b"\x7fELF" or img[4] != 2 or img[5] != 1:
        raise ValueError("not a little-endian ELF64 image")
    shoff, = struct.unpack_from("<Q", img, 0x28)
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", img, 0x3A)
    hdrs = []
    for i in range(shnum):
        name, typ, _flags, _addr, off, size = struct.unpack_from("<IIQQQQ", img, shoff + i * shentsize)
        hdrs.append((name, typ, off, size))
    stroff, strsize = hdrs[shstrndx][2], hdrs[shstrndx][3]
    strtab = img[stroff:stroff + strsize]
    out = []
    for name, typ, off, size in hdrs:
        end = strtab.index(b"\0", name)
        out.append((strtab[name:end].decode(), typ, img[off:off + size]))
    return out


def code_objects(fatbin, arch="gfx950"):
    """The ELF images for `arch` in every bundle of a .hip_fatbin section."""
    out = []
    pos = fatbin.find(MAGIC)
    while pos >= 0:
        n, = struct.unpack_from("<Q", fatbin, pos + len(MAGIC))
        p = pos + len(MAGIC) + 8
        for _ in range(n):
            off, size, tlen = struct.unpack_from("<QQQ", fatbin, p)
            triple = fatbin[p + 24:p + 24 + tlen].decode()
            p += 24 + tlen
            if triple.endswith(arch) and size:
                out.append(fatbin[pos + off:pos + off + size])
        pos = fatbin.find(MAGIC, pos + 1)
    return out


def metadata_notes(img):
    """Decoded NT_AMDGPU_METADATA notes of a code object."""
    out = []
    for _name, typ, data in elf_sections(img):
        if typ != SHT_NOTE:
            continue
        p = 0
        while p + 12 <= len(data):
            namesz, descsz, ntype = struct.unpack_from("<III", data, p)
            p += 12
            owner = data[p:p + namesz].rstrip(b"\0")
            p += (namesz + 3) & ~3
            desc = data[p:p + descsz]
            p += (descsz + 3) & ~3
            if owner == b"AMDGPU" and ntype == NT_AMDGPU_METADATA:
                out.append(msgpack.unpackb(desc, raw=False, strict_map_key=False))
    return out


def kernel_resources(lib_path):
    """{kernel symbol: {field: int}} over every gfx950 kernel of the library."""
    with open(lib_path, "rb") as f:
        lib = f.read()
    fatbins = [data for name, _typ, data in elf_sections(lib) if name == ".hip_fatbin"]
    if not fatbins:
        raise ValueError("no .hip_fatbin section")
    res = {}
    for blob in code_objects(fatbins[0]):
        for md in metadata_notes(blob):
            for k in md.get("amdhsa.kernels", []):
                res[k[".name"]] = {fld: int(k["." + fld]) for fld in FIELDS if "." + fld in k}
    return res
