"""Extracts the golden optimized plans of the reference's Calcite tests as text.

The files algebra/blazingdb-calcite-application/{tpch-reference,selected-queries-reference}.bin
hold java.lang.String objects written by ObjectOutputStream.writeObject
(BlazingRulesTest.java:263-283). They are parsed here as data only: the stream
header (0xACED0005) followed by TC_STRING records (0x74, u16 length,
modified UTF-8 bytes) or TC_LONGSTRING (0x7C, u64 length). No Java object is
instantiated. Output: calcite_plans.txt, one plan per record separated by
blank-line-delimited markers.

Run (in a container that has the reference tree): python tests/golden/extract_plans.py
"""
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/algebra/blazingdb-calcite-application"


def strings(path):
    data = open(path, "rb").read()
    assert data[:4] == b"\xac\xed\x00\x05", "not a Java serialization stream"
    i, out = 4, []
    while i < len(data):
        tag = data[i]
        if tag == 0x74:
            n = struct.unpack(">H", data[i + 1:i + 3])[0]
            out.append(data[i + 3:i + 3 + n].decode("utf-8"))
            i += 3 + n
        elif tag == 0x7C:
            n = struct.unpack(">Q", data[i + 1:i + 9])[0]
            out.append(data[i + 9:i + 9 + n].decode("utf-8"))
            i += 9 + n
        else:
            raise ValueError(f"unexpected record 0x{tag:02x} at {i}")
    return out


def main():
    plans = strings(os.path.join(SRC, "tpch-reference.bin"))
    plans += strings(os.path.join(SRC, "selected-queries-reference.bin"))
    with open(os.path.join(HERE, "calcite_plans.txt"), "w") as f:
        for k, p in enumerate(plans):
            f.write(f"#### plan {k}\n{p.rstrip()}\n")
    print(len(plans), "plans")


if __name__ == "__main__":
    main()
