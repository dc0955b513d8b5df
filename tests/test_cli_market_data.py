"""The hexo CLI's offline market data without a GPU: `hexo -d` (the reference's DOWNLOAD listing,
src/Main.cpp:97-103) over the Tradier JSON responses saved in tests/golden/TRDR (shapes of
src/WebAPI.cpp:53-118: calls only, null volume or ask skipped, a single option object instead of
an array, expiries with T <= 0.1 dropped, working days as src/WebAPI.cpp:22-31 counts them) and
over the compact chain file tests/golden/DEMO.chains."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "hestonexotics_amd", "csrc", "hexo")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(*args, data=GOLD, **env):
    e = dict(os.environ, HEXO_DATA_DIR=data, TZ="UTC", HEXO_TODAY="2024-01-02", **env)
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60, env=e)


def rows(out):
    return [dict(re.findall(r"(\w[\w -]*?): +([-0-9.]+)", l)) for l in out.splitlines() if l.startswith("S:")]


def test_tradier_json_listing():
    out = run("-d", "TRDR")
    assert out.returncode == 0, out.stderr
    r = rows(out.stdout)
    # 2024-01-19 is 13 working days away (T <= 0.1, dropped); 2024-03-15 is 53, 2024-12-20 is 253
    assert [(float(x["strike"]), float(x["ask"]), int(x["volume"]), float(x["expiry time"])) for x in r] == [
        (90.0, 11.7, 400, 53.0), (100.0, 5.0, 900, 53.0), (110.0, 1.7, 250, 53.0), (100.0, 9.1, 600, 253.0)]
    assert all(float(x["S"]) == 100.0 for x in r)


def test_tradier_volume_filter():
    out = run("-d", "TRDR", "volume", "100")
    assert out.returncode == 0, out.stderr
    assert [(float(x["strike"]), int(x["volume"])) for x in rows(out.stdout)] == [(90.0, 120), (100.0, 300)]


def test_chain_file_listing():
    out = run("-d", "DEMO")
    assert out.returncode == 0, out.stderr
    assert len(rows(out.stdout)) == 5


def test_malformed_market_data(tmp_path):
    (tmp_path / "BAD").mkdir()
    (tmp_path / "BAD" / "quote.json").write_text('{"quotes": {"quote": {"bid": 10.0}}}')
    (tmp_path / "BAD" / "expirations.json").write_text('{"expirations": {"date": ["2024-03-15"')
    out = run("-d", "BAD", data=str(tmp_path))
    assert out.returncode == 2 and "JSON" in out.stderr
    out = run("-d", "MISSING", data=str(tmp_path))
    assert out.returncode == 2


@pytest.mark.parametrize("text", [
    "[" * 100_000,                                            # nesting far beyond any API response
    '{"quotes": {"quote": {"bid": "\\uZZZZ"}}}',              # bad unicode escape
    '{"quotes": {"quote": {"bid": 10.0}}} trailing',         # accepted prefix, then junk is ignored
    '{"quotes": {"quote": {"bid": 10.0',                      # truncated
])
def test_hostile_json_is_an_error_not_a_crash(tmp_path, text):
    (tmp_path / "EVIL").mkdir()
    (tmp_path / "EVIL" / "quote.json").write_text(text)
    (tmp_path / "EVIL" / "expirations.json").write_text('{"expirations": {"date": []}}')
    out = run("-d", "EVIL", data=str(tmp_path))
    assert out.returncode in (0, 2), (out.returncode, out.stderr[-300:])
