"""The run-time pair-term composer compiles without a device (hiprtc): a
valid fragment compiles, a broken one reports the compiler's message, and
the functor name must be an identifier."""
from heddlemd_amd import api
from tests import pair_fragments as F


def test_fragments_compile():
    for name, src in (("MorsePair", F.MORSE), ("ZeroPair", F.ZERO)):
        ok, log = api.compile_pair_fragment(name, src)
        assert ok, log


def test_broken_fragment_reports_compiler_log():
    ok, log = api.compile_pair_fragment("MorsePair", F.MORSE.replace("float D", "flot D"))
    assert not ok and "flot" in log


def test_name_must_be_identifier():
    ok, log = api.compile_pair_fragment("Morse Pair", F.MORSE)
    assert not ok and "identifier" in log
