"""Mesh ingestion: the reference's `.tri` and OBJ decoders (host side).

Restates tri_loader (bvhtracer/tri_loader/src/lexer.rs, loader.rs), the
`Mesh` container (bvhtracer/src/mesh/mesh.rs:1-144), `TriMeshDecoder` and
`ObjMeshDecoder` (bvhtracer/src/mesh/decoders.rs:88-215).  Meshes come out
as one (N, 9) float32 array of triangles -- the layout the BVH builder and
the GPU take -- plus per-corner texture coordinates (N, 3, 2) and normals
(N, 3, 3).  Decimal numbers round to binary32 exactly as Rust's
`str::parse::<f32>` does (correctly rounded; a binary64 intermediate is
corrected where it lands on a binary32 midpoint).
"""
import gzip
import io
import os
from fractions import Fraction

import numpy as np

F32 = np.float32


class TriLoaderError(ValueError):
    """tri_loader's TriLoaderError (loader.rs:40-70): line number and kind
    ('ExpectedFloat' or 'EndOfFile')."""

    def __init__(self, line_number, kind, message):
        super().__init__(f"Tri file format parsing error at line {line_number}: {message}")
        self.line_number, self.kind = line_number, kind


def parse_f32(tokens):
    """Decimal strings -> float32, correctly rounded (Rust's f32 parse).
    Raises ValueError on a token Rust would reject."""
    toks = list(tokens)
    for t in toks:
        if "_" in t:
            raise ValueError(f"invalid float literal {t!r}")
    x = np.array([float(t) for t in toks], np.float64)
    with np.errstate(over="ignore"):
        y = x.astype(F32)
    # double rounding can only go wrong where x is exactly a binary32 midpoint
    # (the overflow threshold FLT_MAX + 2^103 included)
    overflow_mid = float(np.finfo(F32).max) + 2.0 ** 103
    for i in np.nonzero(np.isinf(y) & np.isfinite(x) & (np.abs(x) == overflow_mid))[0]:
        if abs(Fraction(toks[i])) < Fraction(overflow_mid):
            y[i] = np.copysign(np.finfo(F32).max, x[i])
    off = (y.astype(np.float64) != x) & np.isfinite(x) & np.isfinite(y)
    for i in np.nonzero(off)[0]:
        toward = F32(np.inf) if x[i] > float(y[i]) else F32(-np.inf)
        with np.errstate(over="ignore"):
            nb = np.nextafter(y[i], toward)
        mid = (float(y[i]) + float(nb)) / 2.0
        if mid != x[i]:
            continue
        exact, m = Fraction(toks[i]), Fraction(mid)
        if exact == m:
            continue  # a true tie: the conversion already rounded it to even
        lo, hi = sorted((y[i], nb), key=float)
        y[i] = hi if exact > m else lo
    return y


def _is_number(tok):
    if tok in ("\n", "\r") or "_" in tok:
        return False
    try:
        float(tok)
    except ValueError:
        return False
    return True


# ------------------------------------------------------------------- .tri

def _tri_tokens(text):
    """tri_loader's lexer (lexer.rs:35-104): whitespace is ' ', '\\\\' and
    '\\t'; each '\\n' or '\\r' is a token of its own; '#' starts a comment to
    the end of the line; anything else runs to the next whitespace or
    newline.  Yields (token, line_number)."""
    i, n, line = 0, len(text), 1
    ws, nl = " \\\t", "\n\r"
    while i < n:
        while i < n and text[i] in ws:
            i += 1
        if i < n and text[i] == "#":
            while i < n and text[i] not in nl:
                i += 1
        if i >= n:
            return
        if text[i] in nl:
            yield text[i], line
            line += 1
            i += 1
            continue
        j = i
        while j < n and text[j] not in ws and text[j] not in nl:
            j += 1
        yield text[i:j], line
        i = j


def load_tri_text(text):
    """TriLoader::load (loader.rs:108-186): triangles of nine numbers,
    separated by '\\n' tokens.  A '\\n' inside a triangle, a '\\r' anywhere a
    number is expected, or a non-number raise TriLoaderError
    ('ExpectedFloat'); input ending inside a triangle raises 'EndOfFile'.
    Returns (N, 9) float32."""
    toks = list(_tri_tokens(text))
    nums, k, line = [], 0, 1
    while k < len(toks):
        tok, ln = toks[k]
        if tok == "\n":
            line += 1
            k += 1
            continue
        for _ in range(9):
            if k >= len(toks):
                raise TriLoaderError(line, "EndOfFile",
                                     "Reached the end of the input in the process of getting the next token.")
            tok, ln = toks[k]
            k += 1
            if tok == "\n":
                line += 1
            if not _is_number(tok):
                raise TriLoaderError(line, "ExpectedFloat",
                                     f"Expected a floating point number but got `{tok}` instead.")
            nums.append(tok)
    return parse_f32(nums).reshape(-1, 9) if nums else np.zeros((0, 9), F32)


def load_tri(path):
    """tri_loader::load(path) (loader.rs); `.gz` files are decompressed."""
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt", newline="") as f:
        return load_tri_text(f.read())


# ------------------------------------------------------------------- Mesh

class Mesh:
    """Mesh<f32> (mesh/mesh.rs): triangles with per-corner texture
    coordinates and normals."""

    def __init__(self, primitives, tex_coords=None, normals=None):
        self._prims = np.ascontiguousarray(primitives, F32).reshape(-1, 9)
        n = self._prims.shape[0]
        self._tex = (np.zeros((n, 3, 2), F32) if tex_coords is None
                     else np.ascontiguousarray(tex_coords, F32).reshape(n, 3, 2))
        self._nrm = (np.zeros((n, 3, 3), F32) if normals is None
                     else np.ascontiguousarray(normals, F32).reshape(n, 3, 3))

    def primitives(self):
        v = self._prims.view()
        v.flags.writeable = False
        return v

    def primitives_mut(self):
        return self._prims

    def tex_coords(self):
        return self._tex

    def normals(self):
        return self._nrm

    def len(self):
        return self._prims.shape[0]

    def len_primitives(self):
        return self._prims.shape[0]

    def permute(self, perm):
        """Reorder the per-triangle attributes like the triangles (the BVH
        build's permutation).  The reference permutes only the triangles
        (bvh.rs:420-429), leaving attributes misaligned."""
        self._tex[:] = self._tex[perm]
        self._nrm[:] = self._nrm[perm]


class MeshBuilder:
    """MeshBuilder (mesh/mesh.rs): with_primitive(triangle, tex_coords, normals)."""

    def __init__(self):
        self.p, self.t, self.n = [], [], []

    def with_primitive(self, triangle, tex_coords=None, normals=None):
        self.p.append(np.asarray(triangle, F32).reshape(9))
        self.t.append(np.zeros((3, 2), F32) if tex_coords is None else np.asarray(tex_coords, F32).reshape(3, 2))
        self.n.append(np.zeros((3, 3), F32) if normals is None else np.asarray(normals, F32).reshape(3, 3))
        return self

    def build(self):
        if not self.p:
            return Mesh(np.zeros((0, 9), F32))
        return Mesh(np.stack(self.p), np.stack(self.t), np.stack(self.n))


def _normalize(v):
    """cglinalg normalize in binary32: v / sqrt((x*x + y*y) + z*z)."""
    d = (v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]
    return v / np.sqrt(d)[..., None]


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


class TriMeshDecoder:
    """TriMeshDecoder (mesh/decoders.rs:88-133): triangles from a .tri
    document; zero texture coordinates; each corner's normal is the face
    normal normalize(normalize(v2 - v0) x normalize(v1 - v0)), in binary32."""

    def __init__(self, reader):
        self.reader = reader

    def read_mesh(self):
        text = self.reader.read()
        if isinstance(text, bytes):
            text = text.decode()
        prims = load_tri_text(text)
        v = prims.reshape(-1, 3, 3)
        with np.errstate(invalid="ignore", divide="ignore"):
            n = _normalize(_cross(_normalize(v[:, 2] - v[:, 0]), _normalize(v[:, 1] - v[:, 0])))
        return Mesh(prims, None, np.repeat(n[:, None, :], 3, axis=1))


# -------------------------------------------------------------------- OBJ

class ObjError(ValueError):
    pass


class ObjMeshDecoder:
    """ObjMeshDecoder (mesh/decoders.rs:136-215) over the OBJ subset the
    reference reads: `v`, `vt`, `vn` and triangular `f` records with
    `v`, `v/vt`, `v//vn` or `v/vt/vn` corners (1-based or negative relative
    indices), faces of the first object only.  Missing texture coordinates
    and normals are zero.  '\\\\' counts as whitespace; other records
    (`g`, `s`, `usemtl`, `mtllib`, `l`, comments) are skipped.  Faces with
    more than three corners are fan-triangulated."""

    def __init__(self, reader):
        self.reader = reader

    def read_mesh(self):
        text = self.reader.read()
        if isinstance(text, bytes):
            text = text.decode()
        v, vt, vn = [], [], []
        tris, texs, nrms = [], [], []
        objects = 0
        for lineno, raw in enumerate(text.splitlines(), 1):
            parts = raw.replace("\\", " ").split("#", 1)[0].split()
            if not parts:
                continue
            tag, args = parts[0], parts[1:]
            if tag == "o":
                objects += 1
                if objects > 1:
                    break
            elif tag == "v":
                v.append(args[:3])
            elif tag == "vt":
                vt.append((args + ["0", "0"])[:2])
            elif tag == "vn":
                vn.append(args[:3])
            elif tag == "f":
                corners = [self._corner(c, len(v), len(vt), len(vn), lineno) for c in args]
                if len(corners) < 3:
                    raise ObjError(f"line {lineno}: a face needs three corners")
                for k in range(1, len(corners) - 1):
                    tri = (corners[0], corners[k], corners[k + 1])
                    tris.append([c[0] for c in tri])
                    texs.append([c[1] for c in tri])
                    nrms.append([c[2] for c in tri])
        vv = parse_f32([x for p in v for x in p]).reshape(-1, 3) if v else np.zeros((0, 3), F32)
        tt = parse_f32([x for p in vt for x in p]).reshape(-1, 2) if vt else np.zeros((0, 2), F32)
        nn = parse_f32([x for p in vn for x in p]).reshape(-1, 3) if vn else np.zeros((0, 3), F32)
        n = len(tris)
        prims = np.zeros((n, 3, 3), F32)
        tex = np.zeros((n, 3, 2), F32)
        nrm = np.zeros((n, 3, 3), F32)
        for i in range(n):
            for k in range(3):
                prims[i, k] = vv[tris[i][k]]
                if texs[i][k] is not None:
                    tex[i, k] = tt[texs[i][k]]
                if nrms[i][k] is not None:
                    nrm[i, k] = nn[nrms[i][k]]
        return Mesh(prims.reshape(n, 9), tex, nrm)

    @staticmethod
    def _corner(tok, nv, nt, nn, lineno):
        fields = tok.split("/")
        if len(fields) > 3 or not fields[0]:
            raise ObjError(f"line {lineno}: bad face corner {tok!r}")

        def index(s, count):
            if s == "":
                return None
            k = int(s)
            k = k - 1 if k > 0 else count + k
            if not 0 <= k < count:
                raise ObjError(f"line {lineno}: index {s} out of range in {tok!r}")
            return k

        vi = index(fields[0], nv)
        ti = index(fields[1], nt) if len(fields) > 1 else None
        ni = index(fields[2], nn) if len(fields) > 2 else None
        return vi, ti, ni


def open_mesh(path):
    """A Mesh from a .tri(.gz) or .obj file."""
    p = str(path)
    opener = gzip.open if p.endswith(".gz") else open
    with opener(p, "rt", newline="") as f:
        data = f.read()
    base = p[:-3] if p.endswith(".gz") else p
    if base.endswith(".obj"):
        return ObjMeshDecoder(io.StringIO(data)).read_mesh()
    if base.endswith(".tri"):
        return TriMeshDecoder(io.StringIO(data)).read_mesh()
    raise ValueError(f"unknown mesh format: {os.path.basename(p)}")
