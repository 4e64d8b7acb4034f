"""Host logic of the Python binding: native objects are released in dependency order (batch, arrays,
BVHs, then the context) whether the caller closes the context or the interpreter exits with everything
still alive. A stand-in for the native library records the calls; no GPU is involved."""
import subprocess
import sys
import textwrap
import pathlib

ROOT = pathlib.Path(__file__).resolve().parents[1]

SCRIPT = textwrap.dedent("""
    import ctypes, os, sys
    sys.path.insert(0, {root!r})
    import vierkant_amd as vk

    handles = [0x1000]
    class Recorder:
        def __getattr__(self, name):
            def call(*args):
                if name in ("vkrt_create", "vkrt_batch_create", "vkrt_device_alloc"):
                    handles[0] += 0x10
                    args[-1]._obj.value = handles[0]
                if "destroy" in name or "free" in name:
                    os.write(1, (name + "\\n").encode())
                return 0
            return call
    vk._lib = Recorder()
    ctx = vk.Context(0)
    bvhs = [vk.Bvh(ctx, ctypes.c_void_p(0x9000 + i)) for i in range(2)]
    rays = [vk.DeviceArray(ctx, 8, vk.RAY_DTYPE) for _ in range(2)]
    hits = [vk.DeviceArray(ctx, 8, vk.HIT_DTYPE) for _ in range(2)]
    batch = vk.Batch(ctx, list(zip(bvhs, rays, hits)))
    if {explicit}:
        ctx.close()
    os.write(1, b"end\\n")
""")


def run(explicit):
    out = subprocess.run([sys.executable, "-c", SCRIPT.format(root=str(ROOT), explicit=explicit)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout.split()


def expected():
    return ["vkrt_batch_destroy"] + ["vkrt_device_free"] * 4 + ["vkrt_bvh_destroy"] * 2 + ["vkrt_destroy"]


def test_release_order_at_interpreter_exit():
    calls = run(False)
    assert calls == ["end"] + expected()


def test_release_order_on_context_close():
    calls = run(True)
    assert calls == expected() + ["end"]
