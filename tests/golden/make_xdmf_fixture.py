"""Writes xdmf3_test_writers.xml: the expected XDMF3 document of the
reference's writer test (test/test_writers.jl:12-34, the XML_ref literal;
grid origin (-0.4, -0.5, 0.0), extent (1.0, 1.1, 1.2), dims (4, 5, 6),
fields Fa and Fb, HDF5 file test.h5).  Run where the reference is checked
out; the tests read only the fixture."""
import os
import sys

src = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "test", "test_writers.jl")
text = open(src).read()
start = text.index('XML_ref = """') + len('XML_ref = """')
end = text.index('"""', start)
xml = text[start:end]
if xml.startswith("\n"):  # a Julia triple-quoted literal drops its first newline
    xml = xml[1:]
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "xdmf3_test_writers.xml")
with open(out, "w") as f:
    f.write(xml)
print("wrote %s (%d bytes)" % (out, len(xml)))
