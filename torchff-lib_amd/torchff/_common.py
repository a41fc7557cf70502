"""Shared guard for the ``use_customized_ops`` switch."""


def require_custom(flag, cls):
    if not flag:
        raise ValueError(f"{cls}: this build provides the HIP implementation only "
                         "(use_customized_ops=True); the pure-PyTorch reference path is not shipped")
