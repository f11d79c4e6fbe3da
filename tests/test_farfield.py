

def test_autotune_key_needs_no_table_lock():
    # the table key is computed where the table lock may already be held
    # (tune saves under it): the device query must not take that lock
    import threading
    from tectosaur_amd import autotune
    assert autotune._ARCH_LOCK is not autotune._LOCK
    done = []
    with autotune._LOCK:
        th = threading.Thread(target=lambda: done.append(autotune._key('elasticH3', 'f32', 64, 64,
                                                                        None)))
        th.start()
        th.join(5.0)
    assert done == ['gfx950|elasticH3|f32|o6|s6']
