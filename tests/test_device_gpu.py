"""Records the device facts the kernels rely on (printed into the test log)."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_device_properties(device):
    p = torch.cuda.get_device_properties(device)
    print(f"[device] {p.name} gcnArch={getattr(p, 'gcnArchName', '?')} CUs={p.multi_processor_count} "
          f"mem={p.total_memory / 2**30:.1f}GiB shared_per_block={getattr(p, 'shared_memory_per_block', '?')} "
          f"shared_per_mp={getattr(p, 'shared_memory_per_multiprocessor', '?')}", flush=True)
    assert "gfx950" in getattr(p, "gcnArchName", "gfx950")
