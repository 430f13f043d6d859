"""Build step: embed the device headers into the runtime library as C strings:
ftr_abi.hpp + ftr_math.hpp + ftr_device.hpp (the source hipRTC specialises per graph) and
ftr_mlp.hpp (the learned proposal's network) and ftr_jtable.hpp (the
J-table), each compiled once."""
import os
import sys

here = os.path.dirname(os.path.abspath(__file__))
abi = open(os.path.join(here, "kernels", "ftr_abi.hpp")).read()
math = open(os.path.join(here, "kernels", "ftr_math.hpp")).read()
dev = open(os.path.join(here, "kernels", "ftr_device.hpp")).read()
dev = dev.replace('#include "ftr_abi.hpp"', "").replace('#include "ftr_math.hpp"', "")
src = (abi + "\n" + math + "\n" + dev).replace("#pragma once", "")
delim = "FTRSRC"
assert (")" + delim) not in src
mlp = open(os.path.join(here, "kernels", "ftr_mlp.hpp")).read().replace("#pragma once", "")
assert (")" + delim) not in mlp
jt = open(os.path.join(here, "kernels", "ftr_jtable.hpp")).read().replace("#pragma once", "")
assert (")" + delim) not in jt
out = sys.argv[1]


def literal(f, name, text):
    # split into chunks: some compilers limit single literal length
    f.write("extern const char* %s;\n" % name)
    f.write("static const char %s_text[] =\n" % name)
    for i in range(0, len(text), 8000):
        f.write('R"%s(%s)%s"\n' % (delim, text[i:i + 8000], delim))
    f.write(";\nconst char* %s = %s_text;\n" % (name, name))


with open(out + ".tmp", "w") as f:
    f.write("/* generated by embed_source.py -- do not edit */\n")
    literal(f, "ftr_embedded_device_source", src)
    literal(f, "ftr_embedded_mlp_source", mlp)
    literal(f, "ftr_embedded_jtable_source", jt)
os.replace(out + ".tmp", out)
