"""The library and torch share one HIP runtime in a process (the library is
loaded after torch whichever module touches it first)."""
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_library_first_then_torch_sees_the_device():
    code = ("import indexing_amd as ia; ia.generate(10);"
            "import torch; assert torch.cuda.is_available();"
            "c = ia.Context(0); print(c.build([3, 1, 3]).keys.tolist()); c.close()")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=300, cwd=__file__.rsplit("/tests/", 1)[0])
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().endswith("[1, 3]")
