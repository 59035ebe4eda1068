"""Device settings (paces/config/_dev_conf.py:29-71), one ROCm GPU per process.

The reference can place lookup tables and vectors on two devices and can back arrays with
managed memory; both exist to stretch memory and are not part of this single-GPU engine
(288 GB of HBM per MI355X). Asking for them raises NotImplementedError.
"""

import torch


class DeviceConfig:
    """Holds the device used by every other object; call ``configure`` first."""

    def __init__(self):
        self.whoami_dev = None
        self.vector_dev = None
        self.mempool = None
        self.split = False
        self.desc_str = ""
        self.configure_called = False

    def _print_and_save(self, s):
        print(s)
        self.desc_str += s + "\n"

    def configure(self, whoami_device_id=0, vector_device_id=0, unified_memory_bytes=None):
        """Select the GPU (paces/config/_dev_conf.py:29-71)."""
        if whoami_device_id != vector_device_id:
            raise NotImplementedError("Splitting lookup tables and vectors over two GPUs is not "
                                      "supported; use one device.")
        if unified_memory_bytes is not None:
            raise NotImplementedError("Managed (unified) memory is not supported.")
        if not torch.cuda.is_available():
            raise RuntimeError("No ROCm GPU is visible to torch; paces_amd needs one.")
        print("######### ROCm device setup: #########")
        self.desc_str = "### ROCm device setup:\n"
        self.whoami_dev = torch.device("cuda", int(whoami_device_id))
        self.vector_dev = self.whoami_dev
        self.split = False
        torch.cuda.set_device(self.vector_dev)
        name = torch.cuda.get_device_name(self.vector_dev)
        self._print_and_save(f"    Using one GPU ({name}).")
        self._print_and_save("    Using torch's caching allocator as the memory pool.")
        self.configure_called = True
        print("#########    end of setup    #########\n")
        self.desc_str += "\n"
