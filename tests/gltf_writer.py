"""Writes glTF 2.0 files for the loader tests: a JSON document plus one binary
buffer, saved as .glb (binary container) or .gltf (buffer as a data: URI or a
file beside it).  Accessors are added from numpy arrays."""
import base64
import json
import struct

import numpy as np

CT = {np.int8: 5120, np.uint8: 5121, np.int16: 5122, np.uint16: 5123, np.uint32: 5125, np.float32: 5126}
TYPE = {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4"}


class Gltf:
    def __init__(self):
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [], "meshes": [],
                    "nodes": [], "materials": [], "textures": [], "images": []}
        self.bin = bytearray()

    def _view(self, data, stride=0):
        while len(self.bin) % 4:
            self.bin.append(0)
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.bin += data
        self.doc["bufferViews"].append(v)
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, arr, normalized=False, minmax=False):
        a = np.ascontiguousarray(arr)
        comps = 1 if a.ndim == 1 else a.shape[1]
        acc = {"bufferView": self._view(a.tobytes()), "componentType": CT[a.dtype.type], "count": int(a.shape[0]),
               "type": TYPE[comps]}
        if normalized:
            acc["normalized"] = True
        if minmax:
            acc["min"] = [float(v) for v in np.atleast_2d(a).min(0)]
            acc["max"] = [float(v) for v in np.atleast_2d(a).max(0)]
        self.doc["accessors"].append(acc)
        return len(self.doc["accessors"]) - 1

    def interleaved(self, arrays):
        """float32 arrays (n, k_i) interleaved in one view -> accessor ids"""
        n = arrays[0].shape[0]
        widths = [a.shape[1] * 4 for a in arrays]
        stride = sum(widths) + 4          # 4 bytes of padding per element
        raw = bytearray(n * stride)
        for i in range(n):
            off = i * stride
            for a, w in zip(arrays, widths):
                raw[off:off + w] = a[i].astype(np.float32).tobytes()
                off += w
        view = self._view(bytes(raw), stride)
        ids, off = [], 0
        for a, w in zip(arrays, widths):
            self.doc["accessors"].append({"bufferView": view, "byteOffset": off, "componentType": 5126,
                                          "count": n, "type": TYPE[a.shape[1]],
                                          "min": [float(v) for v in a.min(0)], "max": [float(v) for v in a.max(0)]})
            ids.append(len(self.doc["accessors"]) - 1)
            off += w
        return ids

    def mesh(self, prims):
        self.doc["meshes"].append({"primitives": prims})
        return len(self.doc["meshes"]) - 1

    def node(self, **kw):
        self.doc["nodes"].append(kw)
        return len(self.doc["nodes"]) - 1

    def material(self, **kw):
        self.doc["materials"].append(kw)
        return len(self.doc["materials"]) - 1

    def image(self, data, embed=True, uri=None):
        if embed:
            self.doc["images"].append({"bufferView": self._view(data), "mimeType": "image/png"})
        else:
            self.doc["images"].append({"uri": uri})
        self.doc["textures"].append({"source": len(self.doc["images"]) - 1})
        return len(self.doc["textures"]) - 1

    def _final(self, buffer_uri=None):
        d = {k: v for k, v in self.doc.items() if v != []}
        buf = {"byteLength": len(self.bin)}
        if buffer_uri:
            buf["uri"] = buffer_uri
        d["buffers"] = [buf]
        return d

    def glb(self):
        js = json.dumps(self._final()).encode()
        js += b" " * (-len(js) % 4)
        b = bytes(self.bin) + bytes(-len(self.bin) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(b), 0x004E4942) + b
        return struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body

    def gltf_data_uri(self):
        uri = "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()
        return json.dumps(self._final(uri)).encode()

    def gltf_external(self, bin_name):
        return json.dumps(self._final(bin_name)).encode(), bytes(self.bin)
