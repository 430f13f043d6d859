"""Build step: embed the device header (ftr_abi.hpp + ftr_device.hpp) into the
runtime library as one C string, the source hipRTC specialises per graph."""
import os
import sys

here = os.path.dirname(os.path.abspath(__file__))
abi = open(os.path.join(here, "kernels", "ftr_abi.hpp")).read()
dev = open(os.path.join(here, "kernels", "ftr_device.hpp")).read()
dev = dev.replace('#include "ftr_abi.hpp"', "")
src = (abi + "\n" + dev).replace("#pragma once", "")
delim = "FTRSRC"
assert (")" + delim) not in src
out = sys.argv[1]
with open(out + ".tmp", "w") as f:
    f.write("/* generated by embed_source.py -- do not edit */\n")
    # split into chunks: some compilers limit single literal length
    f.write("extern const char* ftr_embedded_device_source;\n")
    f.write("static const char ftr_src_text[] =\n")
    for i in range(0, len(src), 8000):
        f.write('R"%s(%s)%s"\n' % (delim, src[i:i + 8000], delim))
    f.write(";\nconst char* ftr_embedded_device_source = ftr_src_text;\n")
os.replace(out + ".tmp", out)
