"""Writes tests/golden/cbor_rfc8949.json: the examples of encoded CBOR data
items of RFC 8949 (Appendix A, Table 6), used to pin the CBOR encoders and
decoders that write and read the reference's files.

Each entry is [kind, value, hex]: kind "uint"/"nint" (value an int), "float"
(value a string for float()), "text", "bytes" (value hex), "bool", "null",
"json" (arrays and string-keyed maps of the above, as JSON), "pairs" (a map
given as a list of [key, value]).  "preferred" is true where the hex is the
preferred (shortest) serialization, which an encoder must reproduce; the
remaining entries (indefinite lengths, tags) are decode-only.  The values are
copied from the RFC, not computed by this repository's code.
"""
import json
import os

PREFERRED = [
    ["uint", 0, "00"], ["uint", 1, "01"], ["uint", 10, "0a"], ["uint", 23, "17"], ["uint", 24, "1818"],
    ["uint", 25, "1819"], ["uint", 100, "1864"], ["uint", 1000, "1903e8"], ["uint", 1000000, "1a000f4240"],
    ["uint", 1000000000000, "1b000000e8d4a51000"], ["uint", 18446744073709551615, "1bffffffffffffffff"],
    ["nint", -1, "20"], ["nint", -10, "29"], ["nint", -100, "3863"], ["nint", -1000, "3903e7"],
    ["float", "0.0", "f90000"], ["float", "-0.0", "f98000"], ["float", "1.0", "f93c00"],
    ["float", "1.1", "fb3ff199999999999a"], ["float", "1.5", "f93e00"], ["float", "65504.0", "f97bff"],
    ["float", "100000.0", "fa47c35000"], ["float", "3.4028234663852886e+38", "fa7f7fffff"],
    ["float", "1.0e+300", "fb7e37e43c8800759c"], ["float", "5.960464477539063e-8", "f90001"],
    ["float", "0.00006103515625", "f90400"], ["float", "-4.0", "f9c400"], ["float", "-4.1", "fbc010666666666666"],
    ["float", "inf", "f97c00"], ["float", "nan", "f97e00"], ["float", "-inf", "f9fc00"],
    ["bool", False, "f4"], ["bool", True, "f5"], ["null", None, "f6"],
    ["bytes", "", "40"], ["bytes", "01020304", "4401020304"],
    ["text", "", "60"], ["text", "a", "6161"], ["text", "IETF", "6449455446"], ["text", "\"\\", "62225c"],
    ["text", "ü", "62c3bc"], ["text", "水", "63e6b0b4"],
    ["json", [], "80"], ["json", [1, 2, 3], "83010203"], ["json", [1, [2, 3], [4, 5]], "8301820203820405"],
    ["json", list(range(1, 26)), "98190102030405060708090a0b0c0d0e0f101112131415161718181819"],
    ["json", {}, "a0"], ["pairs", [[1, 2], [3, 4]], "a201020304"],
    ["json", {"a": 1, "b": [2, 3]}, "a26161016162820203"], ["json", ["a", {"b": "c"}], "826161a161626163"],
    ["json", {"a": "A", "b": "B", "c": "C", "d": "D", "e": "E"}, "a56161614161626142616361436164614461656145"],
]

DECODE_ONLY = [
    ["float", "inf", "fa7f800000"], ["float", "-inf", "faff800000"], ["float", "inf", "fb7ff0000000000000"],
    ["float", "-inf", "fbfff0000000000000"], ["float", "nan", "fa7fc00000"], ["float", "nan", "fb7ff8000000000000"],
    ["uint", 1363896240, "c11a514b67b0"], ["float", "1363896240.5", "c1fb41d452d9ec200000"],
    ["text", "2013-03-21T20:04:00Z", "c074323031332d30332d32315432303a30343a30305a"],
    ["bytes", "0102030405", "5f42010243030405ff"], ["text", "streaming", "7f657374726561646d696e67ff"],
    ["json", [], "9fff"], ["json", [1, [2, 3], [4, 5]], "9f018202039f0405ffff"],
    ["json", [1, [2, 3], [4, 5]], "9f01820203820405ff"], ["json", [1, [2, 3], [4, 5]], "83018202039f0405ff"],
    ["json", [1, [2, 3], [4, 5]], "83019f0203ff820405"],
    ["json", list(range(1, 26)), "9f0102030405060708090a0b0c0d0e0f101112131415161718181819ff"],
    ["json", {"a": 1, "b": [2, 3]}, "bf61610161629f0203ffff"], ["json", ["a", {"b": "c"}], "826161bf61626163ff"],
    ["json", {"Fun": True, "Amt": -2}, "bf6346756ef563416d7421ff"],
]

if __name__ == "__main__":
    data = {"_source": "RFC 8949 Appendix A (examples of encoded CBOR data items); "
                       "tests/golden/make_cbor_vectors.py writes this file.",
            "preferred": PREFERRED, "decode_only": DECODE_ONLY}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cbor_rfc8949.json"), "w") as f:
        json.dump(data, f, indent=0)
        f.write("\n")
